// Reordering of face-1 quadrature points of a connection to match face 0 (Face_permutation.hpp:18-26 and
// Spatial.hpp:73-131 in the hexed tree): a transpose of the rs x rs face for 3D connections between dimensions 0 and 2,
// then a reversal along one tangential axis where the connection flips it.
#ifndef HEXED_AMD_FACE_PERMUTATION_HPP_
#define HEXED_AMD_FACE_PERMUTATION_HPP_

namespace hexed
{

class Face_permutation_dynamic
{
  public:
  virtual ~Face_permutation_dynamic() = default;
  //! reorder face 1 data to match face 0
  virtual void match_faces() = 0;
  //! restore face 1 data to its original order
  virtual void restore() = 0;
};

}  // namespace hexed
#endif
