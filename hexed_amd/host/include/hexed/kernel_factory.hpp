// Kernel base class and the (n_dim, row_size) factory (kernel_factory.hpp:26-119 in the hexed tree).
// The kernels themselves run on the device; n_dim and row_size stay template arguments so that callers of the
// reference's kernel_factory<...>(n_dim, row_size, args...) compile unchanged, and are forwarded at run time.
#ifndef HEXED_AMD_KERNEL_FACTORY_HPP_
#define HEXED_AMD_KERNEL_FACTORY_HPP_

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "Sequence.hpp"
#include "Stopwatch_tree.hpp"

namespace hexed
{

namespace amd
{
void synchronize();  // waits for the device work of the process-wide context
}

constexpr int max_row_size = 8;  // config::max_row_size of the reference build (CMakeLists.txt:74-77)

template <typename T, typename U = void>
class Kernel
{
  // runs the kernel, then waits for its device work, so that a stopwatch running around this call stops after the
  // device finished; an error of the device work is thrown here, never from a destructor
  U timed_call(Sequence<T>& sequence)
  {
    if constexpr (std::is_void_v<U>) {
      (*this)(sequence);
      amd::synchronize();
    } else {
      U result = (*this)(sequence);
      amd::synchronize();
      return result;
    }
  }

  public:
  virtual ~Kernel() = default;
  virtual U operator()(Sequence<T>&) = 0;
  virtual U operator()(Sequence<T>& sequence, Stopwatch_tree& tree)
  {
    tree.work_units_completed += sequence.size();
    Stopwatch::Operator oper(tree.stopwatch);
    return timed_call(sequence);
  }
  virtual U operator()(Sequence<T>& sequence, Stopwatch_tree& category_tree, std::string name)
  {
    auto& specific_tree = category_tree.children.at(name);
    specific_tree.work_units_completed += sequence.size();
    Stopwatch::Operator cat_oper(category_tree.stopwatch);
    Stopwatch::Operator spc_oper(specific_tree.stopwatch);
    return timed_call(sequence);
  }
  typedef std::unique_ptr<Kernel<T, U>> ptr_t;
};

namespace kernel_lookup
{

template <template <int, int> typename kernel>
using ptr_t = typename kernel<1, 2>::ptr_t;

// Every valid (n_dim, row_size) is one index of a flat compile-time range, n_dim-major; a fold over that range
// constructs the one instance whose index matches the requested pair.
constexpr int n_row_sizes = max_row_size - 1;  // row sizes 2 .. max_row_size
constexpr int n_instances = 3 * n_row_sizes;
constexpr int instance_n_dim(int index) { return index / n_row_sizes + 1; }
constexpr int instance_row_size(int index) { return index % n_row_sizes + 2; }

template <template <int, int> typename kernel, int... index, typename... A>
ptr_t<kernel> construct(std::integer_sequence<int, index...>, int n_dim, int row_size, A&... args)
{
  const int wanted = (n_dim - 1) * n_row_sizes + (row_size - 2);
  ptr_t<kernel> made;
  ((index == wanted ? (void)made.reset(new kernel<instance_n_dim(index), instance_row_size(index)>(args...)) : (void)0),
   ...);
  return made;
}

}  // namespace kernel_lookup

//! a freshly constructed kernel<n_dim, row_size>(args...); throws unless 1 <= n_dim <= 3 and 2 <= row_size <= 8
template <template <int, int> typename kernel, typename... A>
kernel_lookup::ptr_t<kernel> kernel_factory(int n_dim, int row_size, A&&... args)
{
  if ((n_dim < 1) || (n_dim > 3) || (row_size < 2) || (row_size > max_row_size)) {
    throw std::runtime_error("demand for invalid kernel");
  }
  return kernel_lookup::construct<kernel>(std::make_integer_sequence<int, kernel_lookup::n_instances>{}, n_dim,
                                          row_size, args...);
}

}  // namespace hexed
#endif
