// Drop-in test: the reference's kernel API (kernel_factory, Spatial<Pde, is_deformed>::..., compute_euler,
// compute_navier_stokes, max_dt_*) driven through hexed_amd's C++ adapter on elements and connections that own their
// storage the way the reference's Element / Face_connection do, compared with the CPU oracle run on an independently
// assembled hx_host_mesh of the same mesh.
//
// usage: test_drop_in [--host-only]   (--host-only runs the checks that need no GPU)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../oracle/hexed_oracle.h"
#include "hexed/Spatial.hpp"
#include "hexed/kernels.hpp"
#include "hexed/pde.hpp"

using namespace hexed;

static int n_fail = 0;
#define REQUIRE(cond)                                                              \
  do {                                                                             \
    if (!(cond)) {                                                                 \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                \
      ++n_fail;                                                                    \
    }                                                                              \
  } while (0)

template <typename F>
static std::string throws(F f)
{
  try {
    f();
  } catch (const std::exception& e) {
    return e.what();
  }
  return "";
}

static int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// ---- storage that mirrors the reference's Element and connections ------------------------------------------------

struct Params
{
  int nd, rs, nq, nfq, nv, face_sz;
  Params(int n_dim, int row_size)
  : nd{n_dim}, rs{row_size}, nq{ipow(row_size, n_dim)}, nfq{ipow(row_size, n_dim - 1)}, nv{n_dim + 2},
    face_sz{std::max(2 * (n_dim + 2), n_dim + row_size) * ipow(row_size, n_dim - 1)}
  {}
};

class Test_element : public Kernel_element
{
  public:
  Params p;
  std::vector<double> data;   // state, time step scale, bulk AV, Laplacian AV, 4 forcing, rs advection (Element.cpp:114-142)
  std::vector<double> rcache;
  std::vector<double> vertex;
  std::vector<double> jac;    // deformed: reference-level normals then determinant
  double h;
  bool def;
  double* faces[6] = {};
  double* fnrml[6] = {};
  double unc = 0;
  Test_element(Params pp, double size, bool deformed)
  : p{pp}, data((pp.nv + 7 + pp.rs) * pp.nq, 0.), rcache(std::max(pp.nv, pp.rs) * pp.nq, 0.), vertex(1 << pp.nd, size / pp.nd), h{size}, def{deformed}
  {
    for (int q = 0; q < p.nq; ++q) data[p.nv * p.nq + q] = 1.;
    if (def) {
      // an undeformed cube: identity reference-level normals, unit determinant, vertex scale h/n_dim
      jac.assign((p.nd * p.nd + 1) * p.nq, 0.);
      for (int i = 0; i < p.nd; ++i)
        for (int q = 0; q < p.nq; ++q) jac[(i * p.nd + i) * p.nq + q] = 1.;
      for (int q = 0; q < p.nq; ++q) jac[p.nd * p.nd * p.nq + q] = 1.;
    }
  }
  double* state() override { return data.data(); }
  double* residual_cache() override { return rcache.data(); }
  double* time_step_scale() override { return data.data() + p.nv * p.nq; }
  double& vertex_time_step_scale(int i) override { return vertex[i]; }
  double nominal_size() override { return h; }
  double* face(int i_face, bool is_ldg) override { return faces[i_face] ? faces[i_face] + is_ldg * p.nv * p.nfq : nullptr; }
  bool deformed() const override { return def; }
  double* reference_level_normals() override { return def ? jac.data() : nullptr; }
  double* jacobian_determinant() override { return def ? jac.data() + p.nd * p.nd * p.nq : nullptr; }
  double* kernel_face_normal(int i_face) override { return def ? fnrml[i_face] : nullptr; }
  double& uncert() override { return unc; }
};

class Test_connection : public Kernel_connection
{
  public:
  Params p;
  Connection_direction dir;
  std::vector<double> data;  // two face blocks, then two normals (connection.hpp:77-86 ordering)
  bool def;
  Test_connection(Params pp, Connection_direction d, bool deformed)
  : p{pp}, dir{d}, data(2 * pp.face_sz + 2 * pp.nd * pp.nfq, 0.), def{deformed}
  {}
  Connection_direction get_direction() override { return dir; }
  double* state(int i_side, bool is_ldg) override { return data.data() + i_side * p.face_sz + is_ldg * p.nv * p.nfq; }
  double* normal() override { return def ? normal_side(0) : nullptr; }
  double* normal_side(int i) { return data.data() + 2 * p.face_sz + i * p.nd * p.nfq; }
};

// a box of n^nd elements on [0, 1]^nd, freestream-like boundary connections on every outer face
struct Box
{
  Params p;
  bool def;
  std::vector<std::unique_ptr<Test_element>> elems;
  std::vector<std::unique_ptr<Test_connection>> cons;      // interior
  std::vector<std::unique_ptr<Test_connection>> bcons;     // boundary (deformed class)
  std::vector<Test_element*> car_v, def_v, all_v;
  std::vector<Test_connection*> car_c, def_c;
  std::vector<Refined_face*> refs;
  std::vector<std::array<int, 2>> con_elems;  // element index per side
  std::vector<int> con_dim;
  std::vector<std::array<int, 3>> bcon_info;  // element, dim, sign
  Box(int nd, int rs, int n, bool deformed) : p{nd, rs}, def{deformed}
  {
    const int ne = ipow(n, nd);
    for (int e = 0; e < ne; ++e) elems.emplace_back(new Test_element(p, 1. / n, def));
    auto idx = [&](int e, int d) { return (e / ipow(n, nd - 1 - d)) % n; };
    for (int d = 0; d < nd; ++d)
      for (int e = 0; e < ne; ++e)
        if (idx(e, d) < n - 1) {
          int f = e + ipow(n, nd - 1 - d);
          Connection_direction dir{{d, d}, {true, false}};
          cons.emplace_back(new Test_connection(p, dir, def));
          auto* c = cons.back().get();
          elems[e]->faces[2 * d + 1] = c->state(0, false);
          elems[f]->faces[2 * d] = c->state(1, false);
          if (def) {
            elems[e]->fnrml[2 * d + 1] = c->normal_side(0);
            elems[f]->fnrml[2 * d] = c->normal_side(1);
            for (int s = 0; s < 2; ++s)
              for (int q = 0; q < p.nfq; ++q) c->normal_side(s)[d * p.nfq + q] = 1.;
          }
          con_elems.push_back({e, f});
          con_dim.push_back(d);
        }
    for (int d = 0; d < nd; ++d)
      for (int s = 0; s < 2; ++s)
        for (int e = 0; e < ne; ++e)
          if (idx(e, d) == (s ? n - 1 : 0)) {
            Connection_direction dir{{d, d}, {bool(s), !bool(s)}};
            bcons.emplace_back(new Test_connection(p, dir, true));
            auto* c = bcons.back().get();
            elems[e]->faces[2 * d + s] = c->state(0, false);
            if (def) elems[e]->fnrml[2 * d + s] = c->normal_side(0);
            for (int side = 0; side < 2; ++side)
              for (int q = 0; q < p.nfq; ++q) c->normal_side(side)[d * p.nfq + q] = 1.;
            bcon_info.push_back({e, d, s});
          }
    for (auto& e : elems) (def ? def_v : car_v).push_back(e.get());
    for (auto& e : elems) all_v.push_back(e.get());
    for (auto& c : cons) (def ? def_c : car_c).push_back(c.get());
    for (auto& c : bcons) def_c.push_back(c.get());
  }
  // smooth seeded state around a freestream, then the faces as Solver::initialize writes them
  void init(double mach, int seed)
  {
    for (size_t e = 0; e < elems.size(); ++e)
      for (int q = 0; q < p.nq; ++q) {
        double* s = elems[e]->state();
        double x = 0.37 * double(e) + 0.11 * q + 0.05 * seed;
        double rho = 1. + 0.05 * std::sin(x), u = mach * (1. + 0.03 * std::cos(1.3 * x));
        double pr = (1. / 1.4) * (1. + 0.04 * std::sin(0.7 * x));
        for (int d = 0; d < p.nd; ++d) s[d * p.nq + q] = d == 0 ? rho * u : 0.01 * rho * std::sin(x + d);
        double ke = 0.;
        for (int d = 0; d < p.nd; ++d) ke += s[d * p.nq + q] * s[d * p.nq + q];
        s[p.nd * p.nq + q] = rho;
        s[(p.nd + 1) * p.nq + q] = pr / 0.4 + 0.5 * ke / rho;
        s[(p.nv + 1) * p.nq + q] = 1e-3;  // bulk AV
      }
  }
};

template <typename T>
struct Vec_seq : public Sequence<T&>
{
  std::vector<T*> v;
  template <typename U>
  explicit Vec_seq(const std::vector<U*>& u) : v(u.begin(), u.end()) {}
  Vec_seq() = default;
  int size() override { return int(v.size()); }
  T& operator[](int i) override { return *v[i]; }
};

// the same mesh assembled directly in the layout of hx_host_mesh, for the oracle (independent of the adapter)
struct Oracle_mesh
{
  Params p;
  std::vector<double> state, rcache, tss, av, nom, vtss, metrics, faces, normals, bc_cache;
  std::vector<int> face_slot, face_nslot, con_slot, con_dir, con_nslot, bc_con, bc_serial;
  std::vector<hx_bc> bcs;
  std::vector<int> car_slot, car_dir, def_slot, def_dir, def_nslot;
  int ne, n_int, n_b;
  bool def;
  explicit Oracle_mesh(Box& b, hx_bc bc) : p{b.p}, def{b.def}
  {
    ne = int(b.elems.size());
    n_int = int(b.cons.size());
    n_b = int(b.bcons.size());
    int n_slot = 2 * (n_int + n_b);
    faces.assign(size_t(n_slot) * p.face_sz, 0.);
    normals.assign(size_t(n_slot) * p.nd * p.nfq, 0.);
    face_slot.assign(ne * 2 * p.nd, 0);
    face_nslot.assign(def ? ne * 2 * p.nd : 0, -1);
    for (int k = 0; k < n_int; ++k) {
      auto ce = b.con_elems[k];
      int d = b.con_dim[k];
      face_slot[ce[0] * 2 * p.nd + 2 * d + 1] = 2 * k;
      face_slot[ce[1] * 2 * p.nd + 2 * d] = 2 * k + 1;
      con_slot.push_back(2 * k);
      con_slot.push_back(2 * k + 1);
      con_dir.push_back(d | (d << 2) | (1 << 4));
      if (def) {
        face_nslot[ce[0] * 2 * p.nd + 2 * d + 1] = 2 * k;
        face_nslot[ce[1] * 2 * p.nd + 2 * d] = 2 * k + 1;
      }
    }
    for (int j = 0; j < n_b; ++j) {
      int k = n_int + j;
      auto bi = b.bcon_info[j];
      face_slot[bi[0] * 2 * p.nd + 2 * bi[1] + bi[2]] = 2 * k;
      if (def) face_nslot[bi[0] * 2 * p.nd + 2 * bi[1] + bi[2]] = 2 * k;
      bc_con.push_back(def ? k : j);
      bc_serial.push_back(0);
    }
    for (int s = 0; s < n_slot; s += 2) {
      int k = s / 2;
      int d = k < n_int ? b.con_dim[k] : b.bcon_info[k - n_int][1];
      for (int side = 0; side < 2; ++side)
        for (int q = 0; q < p.nfq; ++q) normals[(size_t(s + side) * p.nd + d) * p.nfq + q] = 1.;
    }
    bcs.push_back(bc);
    bc_cache.assign(size_t(n_b) * p.nv * p.nfq, 0.);
    for (auto& e : b.elems) {
      state.insert(state.end(), e->data.begin(), e->data.begin() + p.nv * p.nq);
      rcache.insert(rcache.end(), e->rcache.begin(), e->rcache.end());
      tss.insert(tss.end(), e->data.begin() + p.nv * p.nq, e->data.begin() + (p.nv + 1) * p.nq);
      av.insert(av.end(), e->data.begin() + (p.nv + 1) * p.nq, e->data.begin() + (p.nv + 3) * p.nq);
      nom.push_back(e->h);
      vtss.insert(vtss.end(), e->vertex.begin(), e->vertex.end());
      if (def) metrics.insert(metrics.end(), e->jac.begin(), e->jac.end());
    }
  }
  hx_host_mesh desc()
  {
    hx_host_mesh d{};
    d.n_dim = p.nd;
    d.row_size = p.rs;
    int c = def ? 1 : 0;
    d.n_elem[c] = ne;
    d.state[c] = state.data();
    d.rcache[c] = rcache.data();
    d.tss[c] = tss.data();
    d.av[c] = av.data();
    d.nom_size[c] = nom.data();
    d.vertex_tss[c] = vtss.data();
    d.face_slot[c] = face_slot.data();
    d.metrics = def ? metrics.data() : nullptr;
    d.face_nslot = def ? face_nslot.data() : nullptr;
    d.n_slot = int(faces.size() / p.face_sz);
    d.faces = faces.data();
    d.n_nslot = int(normals.size() / (p.nd * p.nfq));
    d.normals = normals.data();
    // interior connections are Cartesian for a Cartesian box; boundary connections are always deformed-class
    std::vector<int>* cs = &con_slot;
    if (def) {
      d.n_con[1] = n_int + n_b;
    } else {
      d.n_con[0] = n_int;
      d.n_con[1] = n_b;
    }
    car_slot.clear();
    car_dir.clear();
    def_slot.clear();
    def_dir.clear();
    def_nslot.clear();
    for (int k = 0; k < n_int; ++k) {
      auto& s = def ? def_slot : car_slot;
      auto& dr = def ? def_dir : car_dir;
      s.push_back((*cs)[2 * k]);
      s.push_back((*cs)[2 * k + 1]);
      dr.push_back(con_dir[k]);
      if (def) {
        def_nslot.push_back(2 * k);
        def_nslot.push_back(2 * k);
      }
    }
    for (int j = 0; j < n_b; ++j) {
      int k = n_int + j;
      def_slot.push_back(2 * k);
      def_slot.push_back(2 * k + 1);
      def_dir.push_back(0);
      def_nslot.push_back(2 * k);
      def_nslot.push_back(2 * k);
    }
    d.con_slot[0] = car_slot.data();
    d.con_dir[0] = car_dir.data();
    d.con_slot[1] = def_slot.data();
    d.con_dir[1] = def_dir.data();
    d.con_nslot = def_nslot.data();
    d.n_bc_con = n_b;
    d.bc_con = bc_con.data();
    d.bc_serial = bc_serial.data();
    d.n_bc = 1;
    d.bcs = bcs.data();
    d.bc_cache = bc_cache.data();
    return d;
  }
};

static double rel_err(const std::vector<double>& a, const std::vector<double>& b)
{
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / den);
}

// face_permutation (kernels_convective.cpp:38-41; Spatial.hpp:73-131) on one face of nd + 2 variables
static void face_permutation_checks()
{
  const int rs = 3, nfq = rs * rs, nv = 5;
  std::vector<double> orig(nv * nfq);
  for (int k = 0; k < nv * nfq; ++k) orig[k] = k;
  // transpose only: dimensions {0, 2} with face signs {1, 1} flip no tangential axis
  std::vector<double> d = orig;
  auto fp = face_permutation(3, rs, Connection_direction{{0, 2}, {true, true}}, d.data());
  fp->match_faces();
  for (int v = 0; v < nv; ++v)
    for (int a = 0; a < rs; ++a)
      for (int b = 0; b < rs; ++b) REQUIRE(d[v * nfq + a * rs + b] == orig[v * nfq + b * rs + a]);
  fp->restore();
  REQUIRE(d == orig);
  // every one of the 3D directions permutes each variable's face points and restores them
  for (int i0 = 0; i0 < 3; ++i0)
    for (int i1 = 0; i1 < 3; ++i1)
      for (int s0 = 0; s0 < 2; ++s0)
        for (int s1 = 0; s1 < 2; ++s1) {
          if (i0 == i1 && s0 == s1) continue;
          d = orig;
          auto p = face_permutation(3, rs, Connection_direction{{i0, i1}, {bool(s0), bool(s1)}}, d.data());
          p->match_faces();
          for (int v = 0; v < nv; ++v) {
            std::vector<double> got(d.begin() + v * nfq, d.begin() + (v + 1) * nfq), want(orig.begin() + v * nfq,
                                                                                          orig.begin() + (v + 1) * nfq);
            std::sort(got.begin(), got.end());
            REQUIRE(got == want);
          }
          p->restore();
          REQUIRE(d == orig);
        }
  // 2D: a flipped tangential axis reverses each variable's row (dimensions {0, 1}, signs {1, 0}: neither normal flips)
  std::vector<double> e(4 * rs);
  for (int k = 0; k < 4 * rs; ++k) e[k] = k;
  face_permutation(2, rs, Connection_direction{{0, 1}, {true, false}}, e.data())->match_faces();
  for (int v = 0; v < 4; ++v)
    for (int i = 0; i < rs; ++i) REQUIRE(e[v * rs + i] == v * rs + rs - 1 - i);
  REQUIRE(throws([&] { face_permutation(4, rs, Connection_direction{{0, 1}, {true, false}}, e.data()); }) ==
          "demand for invalid kernel");
  // the kernel class behind it (kernels_convective.cpp:38-41): the same reordering of Pde::n_extrap variables
  const Connection_direction dir{{2, 0}, {false, false}};
  std::vector<double> a = orig, c = orig;
  face_permutation(3, rs, dir, a.data())->match_faces();
  kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, true>::Face_permutation>(3, rs, dir, c.data())->match_faces();
  REQUIRE(a == c);
  std::vector<double> adv(size_t(3 + rs) * nfq), adv0;
  for (size_t k = 0; k < adv.size(); ++k) adv[k] = double(k);
  adv0 = adv;
  auto fa = kernel_factory<Spatial<pde::Advection, true>::Face_permutation>(3, rs, dir, adv.data());
  fa->match_faces();
  for (int v = 0; v < 3 + rs; ++v)
    for (int i = 0; i < nfq; ++i) REQUIRE(adv[v * nfq + i] - v * nfq == c[i] - 0.);  // every variable alike
  fa->restore();
  REQUIRE(adv == adv0);
}

static void host_only_checks()
{
  face_permutation_checks();
  Gauss_legendre basis(4);
  // kernel_factory range check (kernel_factory.hpp:114-116)
  REQUIRE(throws([&] { kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Local>(4, 4, basis, 1., 0, false, false); }) ==
          "demand for invalid kernel");
  REQUIRE(throws([&] { kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Local>(3, 9, basis, 1., 0, false, false); }) ==
          "demand for invalid kernel");
  // Local's assertions (Spatial.hpp:322-323)
  REQUIRE(throws([&] { kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Local>(3, 4, basis, 1., 1, true, false); }) ==
          "residual calculation is a single-stage operation");
  REQUIRE(throws([&] {
            kernel_factory<Spatial<pde::Navier_stokes<true>::Pde, true>::Local>(3, 4, basis, 1., 1, false, false,
                                                                                Transport_model::constant(1.),
                                                                                Transport_model::constant(1.));
          }) == "two-stage stabilization is not applicable to diffusion equations");
  // basis served from the same tables as the device (Basis.cpp:6-14)
  REQUIRE(std::fabs(basis.step_ratio() - 0.5 / 0.9) < 1e-16);
  REQUIRE(basis.max_cfl() > 0);
  std::printf("host-only checks done\n");
}

// compute_* through the adapter on a registered mesh vs the oracle on the same data
static void compare_update(int nd, int rs, int n, bool def, bool visc, int bc_kind = HX_BC_FREESTREAM)
{
  Box b(nd, rs, n, def);
  b.init(0.5, nd * 10 + rs);
  Gauss_legendre basis(rs);
  Vec_seq<Kernel_element> car(b.car_v), dfm(b.def_v), all(b.all_v);
  Vec_seq<Kernel_connection> car_c(b.car_c), def_c(b.def_c);
  Vec_seq<Refined_face> refs;
  Kernel_mesh km{nd, rs, basis, car_c, def_c, car, dfm, all, refs};
  std::vector<double> fs_state(nd + 2, 0.);
  fs_state[0] = 0.5;
  fs_state[nd] = 1.;
  fs_state[nd + 1] = 1. / 1.4 / 0.4 + 0.125;
  // the reference's constructors with their defaults (Boundary_condition.hpp:66-198)
  hx_bc fs = bc_kind == HX_BC_RIEMANN_INVARIANTS ? amd::riemann_invariants(fs_state)
             : bc_kind == HX_BC_NO_SLIP          ? amd::no_slip_heat_flux()
             : bc_kind == HX_BC_PRESSURE_OUTFLOW ? amd::pressure_outflow(1. / 1.4)
                                                 : amd::freestream(fs_state);
  // initial faces from the element state (Solver::initialize -> compute_write_face)
  {
    amd::Accelerated_mesh am(km);
    for (auto& c : b.bcons) am.add_boundary(*c, fs);
    compute_write_face(km);
    am.pull();
  }
  Oracle_mesh om(b, fs);
  hx_host_mesh od = om.desc();
  // boundary directions {{d, d}, {s, !s}} (connection.hpp:377)
  std::vector<int> dirs(od.con_dir[1], od.con_dir[1] + od.n_con[1]);
  for (int j = 0; j < om.n_b; ++j) {
    auto bi = b.bcon_info[j];
    dirs[(def ? om.n_int : 0) + j] = bi[1] | (bi[1] << 2) | (bi[2] << 4) | ((1 - bi[2]) << 5);
  }
  od.con_dir[1] = dirs.data();
  // the oracle starts from the same faces (copied from the connection buffers the adapter wrote)
  for (int k = 0; k < om.n_int; ++k)
    for (int s = 0; s < 2; ++s)
      std::memcpy(om.faces.data() + size_t(2 * k + s) * b.p.face_sz, b.cons[k]->state(s, false), sizeof(double) * 2 * b.p.nv * b.p.nfq);
  for (int j = 0; j < om.n_b; ++j)
    for (int s = 0; s < 2; ++s)
      std::memcpy(om.faces.data() + size_t(2 * (om.n_int + j) + s) * b.p.face_sz, b.bcons[j]->state(s, false),
                  sizeof(double) * 2 * b.p.nv * b.p.nfq);

  amd::Accelerated_mesh am(km);
  for (auto& c : b.bcons) am.add_boundary(*c, fs);
  Transport_model mu = visc ? Transport_model::constant(2e-3) : inviscid;
  Transport_model kc = visc ? Transport_model::constant(2.8e-3) : inviscid;
  hx_update_options uo{0.5, 1e300, 1, 1, 0, 0, visc, mu.abi(), kc.abi()};
  hx_update_status st{}, so{};
  for (int it = 0; it < 2; ++it) {
    amd::update(am, uo, st);
    REQUIRE(oracle_update(&od, &uo, &so) == 0);
  }
  am.pull();
  std::vector<double> got;
  for (auto& e : b.elems) got.insert(got.end(), e->data.begin(), e->data.begin() + b.p.nv * b.p.nq);
  double err = rel_err(got, om.state);
  std::printf("update nd %d rs %d %s %s bc %d: relative L2 vs oracle %.3e, dt %.6e vs %.6e\n", nd, rs,
              def ? "deformed" : "cartesian", visc ? "navier-stokes" : "euler", bc_kind, err, st.time_step, so.time_step);
  REQUIRE(err <= 1e-12);
  REQUIRE(std::fabs(st.time_step - so.time_step) <= 1e-13 * so.time_step);

  // the kernel classes one by one through kernel_factory, as compute_euler does (kernels_convective.cpp:8-18)
  Stopwatch_tree sw("(element*(time integration stage))");
  sw.children.emplace("neighbor", Stopwatch_tree("(connection*(time integration stage))"));
  sw.children.emplace("local", Stopwatch_tree("(element*(time integration stage))"));
  double dt = 1e-3;
  am.push();
  // after an update the face buffers hold Neighbor's fluxes: write the faces and apply the state conditions first,
  // as a stage does (kernels_convective.cpp:8-18, Solver.cpp:56-67)
  compute_write_face(km);
  amd::check(hx_apply_state_bcs(am.device()), "apply_state_bcs");
  REQUIRE(oracle_write_face(&od, 0, nullptr) == 0);
  REQUIRE(oracle_write_face(&od, 1, nullptr) == 0);
  REQUIRE(oracle_apply_state_bcs(&od) == 0);
  (*kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Neighbor>(nd, rs, 0))(car_c, sw, "neighbor");
  (*kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, true>::Neighbor>(nd, rs, 0))(def_c, sw, "neighbor");
  (*kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Local>(nd, rs, basis, dt, 0, false, false))(car, sw, "local");
  (*kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, true>::Local>(nd, rs, basis, dt, 0, false, false))(dfm, sw, "local");
  am.pull();
  REQUIRE(oracle_neighbor(&od, 0, nullptr, 0) == 0);
  REQUIRE(oracle_neighbor(&od, 1, nullptr, 0) == 0);
  REQUIRE(oracle_local(&od, 0, nullptr, dt, 0, 0, 0) == 0);
  REQUIRE(oracle_local(&od, 1, nullptr, dt, 0, 0, 0) == 0);
  got.clear();
  for (auto& e : b.elems) got.insert(got.end(), e->data.begin(), e->data.begin() + b.p.nv * b.p.nq);
  err = rel_err(got, om.state);
  std::printf("  kernel_factory Neighbor+Local: relative L2 vs oracle %.3e; work units neighbor %d local %d\n", err,
              sw.children.at("neighbor").work_units_completed, sw.children.at("local").work_units_completed);
  REQUIRE(err <= 1e-12);
  REQUIRE(sw.children.at("local").work_units_completed == (long long)b.elems.size());
  // an error inside a timed kernel call reaches the caller as std::runtime_error and leaves the stopwatches stopped
  // (the library's diffusion-only kernels have no compute_residual form: hx_local returns HX_ERR_UNSUPPORTED when
  // the call runs, after the constructor's assertions passed)
  {
    auto bad = kernel_factory<Spatial<pde::Smooth_art_visc, false>::Local>(nd, rs, basis, dt, 0, true, false, 1., 1.);
    const std::string what = throws([&] { (*bad)(car, sw, "local"); });
    std::printf("  timed Local with a failing status: \"%s\"\n", what.c_str());
    REQUIRE(!what.empty());
    REQUIRE(!sw.stopwatch.running() && !sw.children.at("local").stopwatch.running());
  }

  // Solver::is_admissible through the drop-in: admissible, then not after a negative density, then the error text of a
  // non-finite value (thermo.cpp:14)
  am.push();
  REQUIRE(amd::is_admissible(am));
  const double keep = b.elems[0]->data[b.p.nd * b.p.nq + 1];
  b.elems[0]->data[b.p.nd * b.p.nq + 1] = -1.;
  am.push();
  REQUIRE(!amd::is_admissible(am));
  b.elems[0]->data[b.p.nd * b.p.nq + 1] = std::nan("");
  am.push();
  REQUIRE(throws([&] { amd::is_admissible(am); }) == "state is not finite");
  b.elems[0]->data[b.p.nd * b.p.nq + 1] = keep;

  // the admissibility fix's diffusion through the free functions (kernels_max_dt.cpp:21, kernels_diffusive.cpp:32):
  // local time step, ghost states copied, then one step with the device's own flux hook
  am.push();
  Kernel_options fo{sw, sw, sw, 0., 0};
  const double fdt = max_dt_fix_therm_admis(km, fo, 0.7, 0.7, true);
  hx_kernel_options ko{0.4, 0, 0, 0};
  double odt = 0.;
  REQUIRE(oracle_max_dt_fix_therm_admis(&od, &ko, 0.7, 0.7, 1, &odt) == 0);
  REQUIRE(fdt == 1. && odt == 1.);
  amd::check(hx_apply_avc_diff_bcs(am.device()), "apply_avc_diff_bcs");
  REQUIRE(oracle_apply_avc_diff_bcs(&od) == 0);
  fo.dt = 0.4;
  compute_fix_therm_admis(km, fo, {});
  REQUIRE(oracle_compute_fix_therm_admis(&od, &ko) == 0);
  am.pull();
  got.clear();
  for (auto& e : b.elems) got.insert(got.end(), e->data.begin(), e->data.begin() + b.p.nv * b.p.nq);
  err = rel_err(got, om.state);
  std::printf("  compute_fix_therm_admis: relative L2 vs oracle %.3e\n", err);
  REQUIRE(err <= 1e-12);
}

// Max_dt on sequences that were never registered (test_Max_dt.cpp:13-38)
static void unregistered_max_dt()
{
  Params p(1, 2);
  std::vector<std::unique_ptr<Test_element>> els;
  for (int i = 0; i < 2; ++i) els.emplace_back(new Test_element(p, 1.027, false));
  double mass = 0.9, heat_rat = 1.4;
  double e0 = mass * 400. * 400. / (heat_rat * (heat_rat - 1)), e1 = mass * 200. * 200. / (heat_rat * (heat_rat - 1));
  double read[2][6] = {{0, 0, mass, mass, e1, e0}, {0, 0, mass, mass, e0, e1}};
  for (int i = 0; i < 2; ++i) {
    std::memcpy(els[i]->state(), read[i], sizeof(read[i]));
    for (int q = 0; q < 2; ++q) els[i]->time_step_scale()[q] = 0.6;
    els[i]->time_step_scale()[1] = 0.17;
  }
  std::vector<Test_element*> v{els[0].get(), els[1].get()};
  Vec_seq<Kernel_element> view(v);
  Gauss_legendre basis(2);
  double dt = (*kernel_factory<Spatial<pde::Navier_stokes<false>::Pde, false>::Max_dt>(1, 2, basis, false, false, 0.7, 0.7))(view);
  double want = 0.7 * basis.max_cfl() / (400 / 1.027);
  std::printf("unregistered Max_dt: %.15e vs %.15e\n", dt, want);
  REQUIRE(std::fabs(dt - want) <= 1e-13 * want);
  REQUIRE(els[0]->time_step_scale()[1] == 1.);
}

int main(int argc, char** argv)
{
  bool host_only = argc > 1 && std::string(argv[1]) == "--host-only";
  host_only_checks();
  if (!host_only) {
    unregistered_max_dt();
    compare_update(3, 4, 2, false, false);
    compare_update(3, 6, 2, false, true);
    compare_update(2, 5, 3, true, false);
    compare_update(3, 6, 2, true, true);
    compare_update(3, 5, 2, true, true, HX_BC_NO_SLIP);
    compare_update(2, 6, 3, true, true, HX_BC_RIEMANN_INVARIANTS);
    compare_update(3, 4, 2, false, false, HX_BC_PRESSURE_OUTFLOW);
  }
  std::printf("%s: %d failure(s)\n", n_fail ? "FAIL" : "PASS", n_fail);
  return n_fail ? 1 : 0;
}
