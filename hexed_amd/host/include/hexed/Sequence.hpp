// Non-owning sequence view (interface of Sequence.hpp:12-17 in the hexed tree).
#ifndef HEXED_AMD_SEQUENCE_HPP_
#define HEXED_AMD_SEQUENCE_HPP_

#include <vector>

namespace hexed
{

template <typename T>
class Sequence
{
  public:
  virtual ~Sequence() = default;
  virtual int size() = 0;
  virtual T operator[](int index) = 0;
};

//! a Sequence<T&> over a std::vector of pointers (convenience for callers that keep their own objects)
template <typename T, typename Stored = T*>
class Pointer_sequence : public Sequence<T&>
{
  std::vector<Stored>& v;
  public:
  explicit Pointer_sequence(std::vector<Stored>& vec) : v{vec} {}
  int size() override { return int(v.size()); }
  T& operator[](int i) override { return *v[i]; }
};

}  // namespace hexed
#endif
