// Connections as the kernels see them (same members as Kernel_connection.hpp:7-54 in the hexed tree).
#ifndef HEXED_AMD_KERNEL_CONNECTION_HPP_
#define HEXED_AMD_KERNEL_CONNECTION_HPP_

#include <array>

namespace hexed
{

class Connection_direction
{
  public:
  std::array<int, 2> i_dim;
  std::array<bool, 2> face_sign;
  int i_face(int i_side) { return 2 * i_dim[i_side] + face_sign[i_side]; }
  //! does the normal of element `i_side` need flipping to point from element 0 into element 1?
  bool flip_normal(int i_side) { return face_sign[i_side] == bool(i_side); }
  //! does axis i_dim[0] of element 1 need reversing to match element 0?
  bool flip_tangential() { return (i_dim[0] != i_dim[1]) && (flip_normal(0) == flip_normal(1)); }
  //! do the face quadrature points of element 1 need transposing (3D only)?
  bool transpose() { return ((i_dim[0] == 0) && (i_dim[1] == 2)) || ((i_dim[0] == 2) && (i_dim[1] == 0)); }
  //! the packed form of hx_host_mesh::con_dir
  int packed() const { return i_dim[0] | (i_dim[1] << 2) | (int(face_sign[0]) << 4) | (int(face_sign[1]) << 5); }
};

class Connection
{
  public:
  virtual ~Connection() = default;
  virtual Connection_direction get_direction() = 0;
};

class Kernel_connection : virtual public Connection
{
  public:
  //! layout [i_var][i_face_qpoint]
  virtual double* state(int i_side, bool is_ldg) = 0;
  //! area-weighted face normal [i_dim][i_face_qpoint]; nullptr for Cartesian
  virtual double* normal() = 0;
};

}  // namespace hexed
#endif
