// Gauss-Legendre basis (Basis.hpp:14-63, Gauss_legendre.hpp, Basis.cpp:6-14 in the hexed tree), served from the
// tables compiled into libhexed_amd (hx_basis_table), the same tables the device kernels use.
#ifndef HEXED_AMD_BASIS_HPP_
#define HEXED_AMD_BASIS_HPP_

#include <stdexcept>
#include <string>
#include <vector>

#include "../../../../include/hexed_amd.h"

namespace hexed
{

class Basis
{
  protected:
  std::vector<double> table(const char* name) const
  {
    int n = hx_basis_table(row_size, name, nullptr, 0);
    if (n < 0) throw std::runtime_error("Not implemented for required row_size.");
    std::vector<double> v(n);
    hx_basis_table(row_size, name, v.data(), n);
    return v;
  }

  public:
  const int row_size;
  explicit Basis(int row_size_arg) : row_size{row_size_arg}
  {
    if (row_size < HX_MIN_ROW_SIZE || row_size > HX_MAX_ROW_SIZE) throw std::runtime_error("Not implemented for required row_size.");
  }
  virtual ~Basis() = default;
  double node(int i) const { return table("node").at(i); }
  std::vector<double> node_weights() const { return table("weight"); }
  //! row-major [i][j]: derivative of basis polynomial j at node i
  std::vector<double> diff_mat() const { return table("diff_mat"); }
  //! row-major [s][j]: basis polynomial j at s in {0, 1}
  std::vector<double> boundary() const { return table("boundary"); }
  std::vector<double> orthogonal(int degree) const
  {
    auto t = table("orthogonal");
    return std::vector<double>(t.begin() + degree * row_size, t.begin() + (degree + 1) * row_size);
  }
  std::vector<double> filter() const { return table("filter"); }
  std::vector<double> prolong(int i_half) const
  {
    auto t = table("prolong");
    return std::vector<double>(t.begin() + i_half * row_size * row_size, t.begin() + (i_half + 1) * row_size * row_size);
  }
  std::vector<double> restrict(int i_half) const
  {
    auto t = table("restrict");
    return std::vector<double>(t.begin() + i_half * row_size * row_size, t.begin() + (i_half + 1) * row_size * row_size);
  }
  double max_cfl() const { return hx_basis_max_cfl(row_size); }
  double step_ratio() const { return hx_basis_step_ratio(row_size); }
  double min_eig_diffusion() const { return table("eigenvals")[1]; }
};

class Gauss_legendre : public Basis
{
  public:
  explicit Gauss_legendre(int row_size_arg) : Basis(row_size_arg) {}
};

}  // namespace hexed
#endif
