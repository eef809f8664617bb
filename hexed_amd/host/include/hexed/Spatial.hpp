// Spatial<Pde_templ, is_deformed>: the kernel classes of the DG hot path (Spatial.hpp:19-830 in the hexed tree), with
// the same names, constructor arguments and assertions, executed by the gfx950 kernels of libhexed_amd.
#ifndef HEXED_AMD_SPATIAL_HPP_
#define HEXED_AMD_SPATIAL_HPP_

#include <stdexcept>

#include "Basis.hpp"
#include "Kernel_connection.hpp"
#include "Kernel_element.hpp"
#include "Refined_face.hpp"
#include "amd.hpp"
#include "kernel_factory.hpp"
#include "pde.hpp"

namespace hexed
{

template <template <int, int> typename Pde_templ, bool is_deformed>
class Spatial
{
  public:
  Spatial() = delete;

  //! reorders a connection's face-1 data (Pde::n_extrap variables, host memory) to match face 0 (Spatial.hpp:73-131)
  template <int n_dim, int row_size>
  class Face_permutation : public Face_permutation_dynamic
  {
    std::unique_ptr<Face_permutation_dynamic> impl;

    public:
    typedef std::unique_ptr<Face_permutation_dynamic> ptr_t;
    Face_permutation(Connection_direction direction, double* target)
    : impl{row_size < 2 ? nullptr
                        : amd::face_permutation(n_dim, row_size, direction, target, Pde_templ<n_dim, row_size>::n_extrap)}
    {}
    void match_faces() override { impl->match_faces(); }
    void restore() override { impl->restore(); }
  };

  template <int n_dim, int row_size>
  class Write_face : public Kernel<Kernel_element&>
  {
    hx_pde pde;

    public:
    template <typename... pde_args>
    Write_face(const Basis&, pde_args... args) : pde{Pde_templ<n_dim, row_size>(args...).abi}
    {}
    void operator()(Sequence<Kernel_element&>& elements) override
    {
      amd::Mesh_lease m(elements, n_dim, row_size, is_deformed);
      amd::check(hx_write_face(m.device(), m.cls(), &pde), "Write_face");
    }
  };

  template <int n_dim, int row_size>
  class Local : public Kernel<Kernel_element&>
  {
    hx_pde pde;
    double dt;
    int stage;
    bool cr, uf;

    public:
    template <typename... pde_args>
    Local(const Basis&, double dt_arg, bool stage_arg, bool compute_residual, bool use_filter, pde_args... args)
    : pde{Pde_templ<n_dim, row_size>(args...).abi}, dt{dt_arg}, stage{stage_arg}, cr{compute_residual}, uf{use_filter}
    {
      if (Pde_templ<n_dim, row_size>::has_diffusion && stage)
        throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
      if (stage && cr) throw std::runtime_error("residual calculation is a single-stage operation");
    }
    void operator()(Sequence<Kernel_element&>& elements) override
    {
      amd::Mesh_lease m(elements, n_dim, row_size, is_deformed);
      amd::check(hx_local(m.device(), m.cls(), &pde, dt, stage, cr, uf), "Local");
    }
  };

  template <int n_dim, int row_size>
  class Reconcile_ldg_flux : public Kernel<Kernel_element&>
  {
    hx_pde pde;
    double dt;
    int stage;
    bool cr, uf;

    public:
    template <typename... pde_args>
    Reconcile_ldg_flux(const Basis&, double dt_arg, int which_stage, bool compute_residual, bool use_filter,
                       pde_args... args)
    : pde{Pde_templ<n_dim, row_size>(args...).abi}, dt{dt_arg}, stage{which_stage}, cr{compute_residual}, uf{use_filter}
    {
      if (Pde_templ<n_dim, row_size>::has_diffusion && stage)
        throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    }
    void operator()(Sequence<Kernel_element&>& elements) override
    {
      amd::Mesh_lease m(elements, n_dim, row_size, is_deformed);
      amd::check(hx_reconcile_ldg_flux(m.device(), m.cls(), &pde, dt, stage, cr, uf), "Reconcile_ldg_flux");
    }
  };

  template <int n_dim, int row_size>
  class Neighbor : public Kernel<Kernel_connection&>
  {
    hx_pde pde;
    int stage;

    public:
    template <typename... pde_args>
    Neighbor(int i_stage, pde_args... args) : pde{Pde_templ<n_dim, row_size>(args...).abi}, stage{i_stage}
    {}
    void operator()(Sequence<Kernel_connection&>& connections) override
    {
      amd::Mesh_lease m(connections, n_dim, row_size, is_deformed);
      amd::check(hx_neighbor(m.device(), m.cls(), &pde, stage), "Neighbor");
    }
  };

  template <int n_dim, int row_size>
  class Neighbor_reconcile : public Kernel<Kernel_connection&>
  {
    hx_pde pde;

    public:
    Neighbor_reconcile() : pde{pde::kind_of<Pde_templ<n_dim, row_size>>()} {}
    void operator()(Sequence<Kernel_connection&>& connections) override
    {
      amd::Mesh_lease m(connections, n_dim, row_size, is_deformed);
      amd::check(hx_neighbor_reconcile(m.device(), m.cls(), &pde), "Neighbor_reconcile");
    }
  };

  template <int n_dim, int row_size>
  class Max_dt : public Kernel<Kernel_element&, double>
  {
    hx_pde pde;
    bool local, uf;
    double sc, sd;

    public:
    template <typename... pde_args>
    Max_dt(const Basis&, bool is_local, bool use_filter, double safety_conv, double safety_diff, pde_args... args)
    : pde{Pde_templ<n_dim, row_size>(args...).abi}, local{is_local}, uf{use_filter}, sc{safety_conv}, sd{safety_diff}
    {}
    double operator()(Sequence<Kernel_element&>& elements) override
    {
      amd::Mesh_lease m(elements, n_dim, row_size, is_deformed);
      double dt = 0.;
      amd::check(hx_max_dt(m.device(), m.cls(), &pde, local, uf, sc, sd, &dt), "Max_dt");
      return dt;
    }
  };

  template <int n_dim, int row_size>
  class Prolong_refined : public Kernel<Refined_face&>
  {
    bool scl, off;

    public:
    Prolong_refined(const Basis&, bool scale = false, bool offset = false) : scl{scale}, off{offset} {}
    void operator()(Sequence<Refined_face&>& ref_faces) override
    {
      amd::Mesh_lease m(ref_faces, n_dim, row_size);
      hx_pde pde = pde::kind_of<Pde_templ<n_dim, row_size>>();
      amd::check(hx_prolong_refined(m.device(), &pde, scl, off), "Prolong_refined");
    }
  };

  template <int n_dim, int row_size>
  class Restrict_refined : public Kernel<Refined_face&>
  {
    bool scl, off;

    public:
    Restrict_refined(const Basis&, bool scale = true, bool offset = false) : scl{scale}, off{offset} {}
    void operator()(Sequence<Refined_face&>& ref_faces) override
    {
      amd::Mesh_lease m(ref_faces, n_dim, row_size);
      hx_pde pde = pde::kind_of<Pde_templ<n_dim, row_size>>();
      amd::check(hx_restrict_refined(m.device(), &pde, scl, off), "Restrict_refined");
    }
  };
};

}  // namespace hexed
#endif
