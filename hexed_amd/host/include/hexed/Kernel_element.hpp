// The element as the kernels see it (same pure-virtual interface as Kernel_element.hpp:14-45 in the hexed tree).
// Storage order is the reference's: state()[i_var*n_qpoint + i_qpoint] for the n_dim + 2 conserved variables,
// followed in the same block by the time step scale (offset n_dim + 2) and the bulk / Laplacian artificial
// viscosity coefficients (offsets n_dim + 3, n_dim + 4), as pde.hpp:17-21 lays them out.
#ifndef HEXED_AMD_KERNEL_ELEMENT_HPP_
#define HEXED_AMD_KERNEL_ELEMENT_HPP_

namespace hexed
{

class Kernel_element
{
  public:
  virtual ~Kernel_element() = default;
  virtual double* state() = 0;
  virtual double* residual_cache() = 0;
  virtual double* time_step_scale() = 0;
  virtual double& vertex_time_step_scale(int i_vertex) = 0;
  virtual double nominal_size() = 0;
  virtual double* face(int i_face, bool is_ldg) = 0;
  virtual bool deformed() const = 0;
  virtual double* reference_level_normals() = 0;   // [i_dim][j_dim][i_qpoint]; nullptr for Cartesian
  virtual double* jacobian_determinant() = 0;      // [i_qpoint]; nullptr for Cartesian
  virtual double* kernel_face_normal(int i_face) = 0;  // [i_dim][i_face_qpoint]; nullptr => Cartesian neighbour
  virtual double& uncert() = 0;
};

}  // namespace hexed
#endif
