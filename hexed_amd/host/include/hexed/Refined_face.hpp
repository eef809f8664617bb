// Hanging-face bookkeeping (Refined_face.hpp:9-15 in the hexed tree).
#ifndef HEXED_AMD_REFINED_FACE_HPP_
#define HEXED_AMD_REFINED_FACE_HPP_

#include <array>

namespace hexed
{

class Refined_face
{
  public:
  double* coarse = nullptr;
  std::array<double*, 4> fine{};
  std::array<bool, 2> stretch{false, false};
};

}  // namespace hexed
#endif
