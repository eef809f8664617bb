// Kernel base class and the (n_dim, row_size) factory (kernel_factory.hpp:26-119 in the hexed tree).
// The kernels themselves run on the device; n_dim and row_size stay template arguments so that callers of the
// reference's kernel_factory<...>(n_dim, row_size, args...) compile unchanged, and are forwarded at run time.
#ifndef HEXED_AMD_KERNEL_FACTORY_HPP_
#define HEXED_AMD_KERNEL_FACTORY_HPP_

#include <memory>
#include <stdexcept>
#include <string>

#include "Sequence.hpp"
#include "Stopwatch_tree.hpp"

namespace hexed
{

namespace amd
{
void synchronize();  // waits for the device work of the process-wide context
}

constexpr int max_row_size = 8;  // config::max_row_size of the reference build (CMakeLists.txt:74-77)

//! waits for the device in its destructor, so a stopwatch declared before it stops after the device work finished
struct Device_sync_guard
{
  ~Device_sync_guard() { amd::synchronize(); }
};

template <typename T, typename U = void>
class Kernel
{
  public:
  virtual ~Kernel() = default;
  virtual U operator()(Sequence<T>&) = 0;
  virtual U operator()(Sequence<T>& sequence, Stopwatch_tree& tree)
  {
    tree.work_units_completed += sequence.size();
    Stopwatch::Operator oper(tree.stopwatch);
    Device_sync_guard sync;
    return (*this)(sequence);
  }
  virtual U operator()(Sequence<T>& sequence, Stopwatch_tree& category_tree, std::string name)
  {
    auto& specific_tree = category_tree.children.at(name);
    specific_tree.work_units_completed += sequence.size();
    Stopwatch::Operator cat_oper(category_tree.stopwatch);
    Stopwatch::Operator spc_oper(specific_tree.stopwatch);
    Device_sync_guard sync;
    return (*this)(sequence);
  }
  typedef std::unique_ptr<Kernel<T, U>> ptr_t;
};

namespace kernel_lookup
{

template <template <int, int> typename kernel>
using ptr_t = typename kernel<1, 2>::ptr_t;

// compile-time list of every (n_dim, row_size) instance, searched from (max_n_dim, max_row_size) downwards
template <template <int, int> typename kernel, int max_n_dim, int max_row_size_>
class Kernel_lookup
{
  Kernel_lookup<kernel, max_n_dim, max_row_size_ - 1> next;

  public:
  template <typename... A>
  ptr_t<kernel> get(int n_dim, int row_size, A&&... args)
  {
    if (n_dim == max_n_dim && row_size == max_row_size_) return ptr_t<kernel>{new kernel<max_n_dim, max_row_size_>(args...)};
    return next.get(n_dim, row_size, args...);
  }
};

template <template <int, int> typename kernel, int max_n_dim>
class Kernel_lookup<kernel, max_n_dim, 1>
{
  Kernel_lookup<kernel, max_n_dim - 1, max_row_size> next;

  public:
  template <typename... A>
  ptr_t<kernel> get(int n_dim, int row_size, A&&... args)
  {
    if (n_dim == max_n_dim) return ptr_t<kernel>{new kernel<max_n_dim, 1>(args...)};
    return next.get(n_dim, row_size, args...);
  }
};

template <template <int, int> typename kernel>
class Kernel_lookup<kernel, 1, 1>
{
  public:
  template <typename... A>
  ptr_t<kernel> get(int, int, A&&... args)
  {
    return ptr_t<kernel>{new kernel<1, 1>(args...)};
  }
};

}  // namespace kernel_lookup

//! a freshly constructed kernel<n_dim, row_size>(args...); throws unless 1 <= n_dim <= 3 and 2 <= row_size <= 8
template <template <int, int> typename kernel, typename... A>
kernel_lookup::ptr_t<kernel> kernel_factory(int n_dim, int row_size, A&&... args)
{
  if ((n_dim < 1) || (n_dim > 3) || (row_size < 2) || (row_size > max_row_size)) {
    throw std::runtime_error("demand for invalid kernel");
  }
  kernel_lookup::Kernel_lookup<kernel, 3, max_row_size> lookup;
  return lookup.get(n_dim, row_size, args...);
}

}  // namespace hexed
#endif
