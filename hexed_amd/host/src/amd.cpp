// Device registry of the drop-in (hexed/amd.hpp) and the free functions of hexed/kernels.hpp.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "hexed/Spatial.hpp"
#include "hexed/amd.hpp"
#include "hexed/kernels.hpp"

namespace hexed::amd
{

namespace
{

std::mutex registry_mutex;
std::unordered_map<const void*, std::pair<Accelerated_mesh*, int>>& registry()
{
  static std::unordered_map<const void*, std::pair<Accelerated_mesh*, int>> r;
  return r;
}

int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

template <typename T>
std::vector<T*> collect(Sequence<T&>& s)
{
  std::vector<T*> out;
  out.reserve(s.size());
  for (int i = 0; i < s.size(); ++i) out.push_back(&s[i]);
  return out;
}

}  // namespace

hx_ctx* context()
{
  static hx_ctx* ctx = [] {
    hx_ctx* c = nullptr;
    const char* dev = std::getenv("HEXED_AMD_DEVICE");
    hx_status s = hx_ctx_create(dev ? std::atoi(dev) : 0, &c);
    if (s != HX_OK) throw std::runtime_error("hexed_amd: no usable GPU (hx_ctx_create status " + std::to_string(s) + ")");
    return c;
  }();
  return ctx;
}

void synchronize() { check(hx_ctx_synchronize(context()), "synchronize"); }

void check(hx_status s, const char* what)
{
  if (s == HX_OK) return;
  // the reference's own messages (kernel_factory, HEXED_ASSERT) come back verbatim
  throw std::runtime_error(std::string(hx_last_error(context())).empty() ? std::string(what)
                                                                        : std::string(hx_last_error(context())));
}

Accelerated_mesh::Accelerated_mesh(Kernel_mesh mesh, bool reference_layout)
: Accelerated_mesh(mesh.n_dim, mesh.row_size, collect(mesh.car_elems), collect(mesh.def_elems),
                   collect(mesh.car_cons), collect(mesh.def_cons), collect(mesh.ref_faces), reference_layout, false)
{
  std::lock_guard<std::mutex> lock(registry_mutex);
  keys = {&mesh.car_elems, &mesh.def_elems, &mesh.car_cons, &mesh.def_cons, &mesh.ref_faces, &mesh.elems};
  registry()[&mesh.car_elems] = {this, HX_CARTESIAN};
  registry()[&mesh.def_elems] = {this, HX_DEFORMED};
  registry()[&mesh.car_cons] = {this, HX_CARTESIAN};
  registry()[&mesh.def_cons] = {this, HX_DEFORMED};
  registry()[&mesh.ref_faces] = {this, 0};
  registry()[&mesh.elems] = {this, -1};
}

Accelerated_mesh::Accelerated_mesh(int n_dim, int row_size, std::vector<Kernel_element*> car,
                                   std::vector<Kernel_element*> def, std::vector<Kernel_connection*> car_cons,
                                   std::vector<Kernel_connection*> def_cons, std::vector<Refined_face*> ref_faces,
                                   bool reference_layout, bool)
: nd{n_dim}, rs{row_size}, ref_layout{reference_layout}, refs{std::move(ref_faces)}
{
  if (nd < 1 || nd > 3 || rs < 2 || rs > max_row_size) throw std::runtime_error("demand for invalid kernel");
  nq = ipow(rs, nd);
  nfq = nq / rs;
  nv = nd + 2;
  face_sz = std::max(2 * nv, nd + rs) * nfq;
  elems[0] = std::move(car);
  elems[1] = std::move(def);
  cons[0] = std::move(car_cons);
  cons[1] = std::move(def_cons);
  build_topology();
}

Accelerated_mesh::~Accelerated_mesh()
{
  {
    std::lock_guard<std::mutex> lock(registry_mutex);
    for (const void* k : keys) {
      auto it = registry().find(k);
      if (it != registry().end() && it->second.first == this) registry().erase(it);
    }
  }
  if (dev) hx_mesh_destroy(dev);
}

Accelerated_mesh* Accelerated_mesh::find(const void* sequence, int* cls)
{
  std::lock_guard<std::mutex> lock(registry_mutex);
  auto it = registry().find(sequence);
  if (it == registry().end()) return nullptr;
  if (cls) *cls = it->second.second;
  return it->second.first;
}

int Accelerated_mesh::slot(double* p)
{
  auto it = slot_of.find(p);
  if (it != slot_of.end()) return it->second;
  int s = int(slot_ptr.size());
  slot_ptr.push_back(p);
  if (p) slot_of[p] = s;
  return s;
}

int Accelerated_mesh::nslot(double* p)
{
  if (!p) return -1;
  auto it = nslot_of.find(p);
  if (it != nslot_of.end()) return it->second;
  int s = int(nslot_ptr.size());
  nslot_ptr.push_back(p);
  nslot_of[p] = s;
  return s;
}

void Accelerated_mesh::build_topology()
{
  int dummy = -1;
  for (int c = 0; c < 2; ++c) {
    face_slot[c].assign(elems[c].size() * 2 * nd, 0);
    for (size_t e = 0; e < elems[c].size(); ++e) {
      Kernel_element& el = *elems[c][e];
      for (int f = 0; f < 2 * nd; ++f) {
        double* p = el.face(f, false);
        if (!p) {
          if (dummy < 0) dummy = slot(nullptr);
          face_slot[c][e * 2 * nd + f] = dummy;
          continue;
        }
        if (el.face(f, true) != p + nv * nfq)
          throw std::runtime_error("hexed_amd: face storage must hold the LDG block right after the state block");
        face_slot[c][e * 2 * nd + f] = slot(p);
      }
    }
  }
  face_nslot.assign(elems[1].size() * 2 * nd, -1);
  for (size_t e = 0; e < elems[1].size(); ++e)
    for (int f = 0; f < 2 * nd; ++f) face_nslot[e * 2 * nd + f] = nslot(elems[1][e]->kernel_face_normal(f));
  for (int c = 0; c < 2; ++c) {
    con_slot[c].clear();
    con_dir[c].clear();
    for (Kernel_connection* con : cons[c]) {
      con_slot[c].push_back(slot(con->state(0, false)));
      con_slot[c].push_back(slot(con->state(1, false)));
      Connection_direction d = con->get_direction();
      con_dir[c].push_back(c == HX_CARTESIAN ? Connection_direction{{d.i_dim[0], d.i_dim[0]}, {true, false}}.packed()
                                             : d.packed());
      if (c == HX_DEFORMED) {
        double* n = con->normal();
        if (!n) throw std::runtime_error("hexed_amd: deformed connection without a normal");
        int ns = nslot(n);
        con_nslot.push_back(ns);
        con_nslot.push_back(ns);
      }
    }
  }
  const int n_face = 1 << (nd - 1);
  for (Refined_face* r : refs) {
    ref_slot.push_back(slot(r->coarse));
    int nf = n_face;
    for (int d = 0; d < nd - 1; ++d) nf /= 1 + r->stretch[d];
    for (int k = 0; k < 4; ++k) ref_slot.push_back(k < nf ? slot(r->fine[k]) : -1);
    ref_stretch.push_back(r->stretch[0]);
    ref_stretch.push_back(r->stretch[1]);
  }
  for (int c = 0; c < 2; ++c) {
    const size_t n = elems[c].size();
    state[c].assign(n * nv * nq, 0.);
    rcache[c].assign(n * std::max(nv, rs) * nq, 0.);
    tss[c].assign(n * nq, 1.);
    av[c].assign(ref_layout ? n * 2 * nq : 0, 0.);
    forcing[c].assign(ref_layout ? n * 4 * nq : 0, 0.);
    advection[c].assign(ref_layout ? n * rs * nq : 0, 0.);
    nom[c].resize(n);
    vtss[c].resize(n << nd);
    for (size_t e = 0; e < n; ++e) {
      nom[c][e] = elems[c][e]->nominal_size();
      for (int v = 0; v < (1 << nd); ++v) vtss[c][(e << nd) + v] = elems[c][e]->vertex_time_step_scale(v);
    }
  }
  metrics.assign(elems[1].size() * (nd * nd + 1) * nq, 0.);
  faces.assign(slot_ptr.size() * face_sz, 0.);
  normals.assign(nslot_ptr.size() * nd * nfq, 0.);
}

void Accelerated_mesh::add_boundary(Kernel_connection& con, hx_bc bc)
{
  if (dev) throw std::runtime_error("hexed_amd: boundary conditions must be added before the first push");
  auto it = std::find(cons[1].begin(), cons[1].end(), &con);
  if (it == cons[1].end()) throw std::runtime_error("hexed_amd: boundary connection is not in def_cons");
  // field by field: hx_bc has padding after its int members, which memcmp would compare too
  auto same = [](const hx_bc& a, const hx_bc& b) {
    return a.kind == b.kind && std::equal(a.state, a.state + HX_MAX_N_DIM + 2, b.state) && a.pressure == b.pressure &&
           a.thermal == b.thermal && a.thermal_value == b.thermal_value && a.emissivity == b.emissivity &&
           a.heat_transfer_coef == b.heat_transfer_coef && a.temperature == b.temperature && a.coercion == b.coercion;
  };
  int serial = int(std::find_if(bc_defs.begin(), bc_defs.end(), [&](const hx_bc& b) { return same(b, bc); }) -
                   bc_defs.begin());
  if (serial == int(bc_defs.size())) bc_defs.push_back(bc);
  bc_con.push_back(int(it - cons[1].begin()));
  bc_serial.push_back(serial);
}

void Accelerated_mesh::gather()
{
  for (int c = 0; c < 2; ++c) {
    for (size_t e = 0; e < elems[c].size(); ++e) {
      Kernel_element& el = *elems[c][e];
      const double* st = el.state();
      std::copy(st, st + nv * nq, state[c].begin() + e * nv * nq);
      const double* rc = el.residual_cache();
      const size_t nrc = std::max(nv, rs) * nq;  // residual cache: max(nv, rs) variables (Storage_params.cpp:32-35)
      std::copy(rc, rc + nrc, rcache[c].begin() + e * nrc);
      const double* ts = el.time_step_scale();
      std::copy(ts, ts + nq, tss[c].begin() + e * nq);
      if (ref_layout) {
        std::copy(st + (nv + 1) * nq, st + (nv + 3) * nq, av[c].begin() + e * 2 * nq);
        std::copy(st + (nv + 3) * nq, st + (nv + 7) * nq, forcing[c].begin() + e * 4 * nq);
        std::copy(st + (nv + 7) * nq, st + (nv + 7 + rs) * nq, advection[c].begin() + e * rs * nq);
      }
      if (c == HX_DEFORMED) {
        const double* n = el.reference_level_normals();
        const double* d = el.jacobian_determinant();
        std::copy(n, n + nd * nd * nq, metrics.begin() + e * (nd * nd + 1) * nq);
        std::copy(d, d + nq, metrics.begin() + e * (nd * nd + 1) * nq + nd * nd * nq);
      }
    }
  }
  for (size_t s = 0; s < slot_ptr.size(); ++s)
    if (slot_ptr[s]) std::copy(slot_ptr[s], slot_ptr[s] + face_sz, faces.begin() + s * face_sz);
  for (size_t s = 0; s < nslot_ptr.size(); ++s)
    std::copy(nslot_ptr[s], nslot_ptr[s] + nd * nfq, normals.begin() + s * nd * nfq);
}

void Accelerated_mesh::scatter()
{
  for (int c = 0; c < 2; ++c) {
    for (size_t e = 0; e < elems[c].size(); ++e) {
      Kernel_element& el = *elems[c][e];
      std::copy(state[c].begin() + e * nv * nq, state[c].begin() + (e + 1) * nv * nq, el.state());
      const size_t nrc = std::max(nv, rs) * nq;
      std::copy(rcache[c].begin() + e * nrc, rcache[c].begin() + (e + 1) * nrc, el.residual_cache());
      std::copy(tss[c].begin() + e * nq, tss[c].begin() + (e + 1) * nq, el.time_step_scale());
      if (ref_layout) {
        double* st = el.state();
        std::copy(av[c].begin() + e * 2 * nq, av[c].begin() + (e + 1) * 2 * nq, st + (nv + 1) * nq);
        std::copy(forcing[c].begin() + e * 4 * nq, forcing[c].begin() + (e + 1) * 4 * nq, st + (nv + 3) * nq);
        std::copy(advection[c].begin() + e * rs * nq, advection[c].begin() + (e + 1) * rs * nq, st + (nv + 7) * nq);
      }
    }
  }
  for (size_t s = 0; s < slot_ptr.size(); ++s)
    if (slot_ptr[s]) std::copy(faces.begin() + s * face_sz, faces.begin() + (s + 1) * face_sz, slot_ptr[s]);
}

hx_host_mesh Accelerated_mesh::description()
{
  hx_host_mesh d{};
  d.n_dim = nd;
  d.row_size = rs;
  for (int c = 0; c < 2; ++c) {
    d.n_elem[c] = int(elems[c].size());
    d.state[c] = state[c].data();
    d.rcache[c] = rcache[c].data();
    d.tss[c] = tss[c].data();
    d.av[c] = av[c].empty() ? nullptr : av[c].data();
    d.forcing[c] = forcing[c].empty() ? nullptr : forcing[c].data();
    d.advection[c] = advection[c].empty() ? nullptr : advection[c].data();
    d.nom_size[c] = nom[c].data();
    d.vertex_tss[c] = vtss[c].data();
    d.face_slot[c] = face_slot[c].data();
    d.n_con[c] = int(cons[c].size());
    d.con_slot[c] = con_slot[c].data();
    d.con_dir[c] = con_dir[c].data();
  }
  d.metrics = metrics.data();
  d.face_nslot = face_nslot.data();
  d.n_slot = int(slot_ptr.size());
  d.faces = faces.data();
  d.n_nslot = int(nslot_ptr.size());
  d.normals = normals.data();
  d.con_nslot = con_nslot.data();
  d.n_ref = int(refs.size());
  d.ref_slot = ref_slot.data();
  d.ref_stretch = ref_stretch.data();
  d.n_bc_con = int(bc_con.size());
  d.bc_con = bc_con.data();
  d.bc_serial = bc_serial.data();
  d.n_bc = int(bc_defs.size());
  d.bcs = bc_defs.data();
  return d;
}

hx_mesh* Accelerated_mesh::device()
{
  if (!dev) push();
  return dev;
}

void Accelerated_mesh::push()
{
  gather();
  if (!dev) {
    hx_host_mesh d = description();
    check(hx_mesh_create(context(), &d, &dev), "hx_mesh_create");
    return;
  }
  for (int c = 0; c < 2; ++c) {
    if (elems[c].empty()) continue;
    check(hx_mesh_upload(dev, HX_FIELD_STATE, c, state[c].data(), state[c].size()), "upload");
    check(hx_mesh_upload(dev, HX_FIELD_RCACHE, c, rcache[c].data(), rcache[c].size()), "upload");
    check(hx_mesh_upload(dev, HX_FIELD_TSS, c, tss[c].data(), tss[c].size()), "upload");
    if (!av[c].empty()) check(hx_mesh_upload(dev, HX_FIELD_AV, c, av[c].data(), av[c].size()), "upload");
    if (!forcing[c].empty()) check(hx_mesh_upload(dev, HX_FIELD_FORCING, c, forcing[c].data(), forcing[c].size()), "upload");
    if (!advection[c].empty())
      check(hx_mesh_upload(dev, HX_FIELD_ADVECTION, c, advection[c].data(), advection[c].size()), "upload");
  }
  if (!elems[1].empty()) check(hx_mesh_upload(dev, HX_FIELD_METRICS, 1, metrics.data(), metrics.size()), "upload");
  check(hx_mesh_upload(dev, HX_FIELD_FACES, 0, faces.data(), faces.size()), "upload");
  if (!normals.empty()) check(hx_mesh_upload(dev, HX_FIELD_NORMALS, 0, normals.data(), normals.size()), "upload");
}

void Accelerated_mesh::pull()
{
  if (!dev) return;
  for (int c = 0; c < 2; ++c) {
    if (elems[c].empty()) continue;
    check(hx_mesh_download(dev, HX_FIELD_STATE, c, state[c].data(), state[c].size()), "download");
    check(hx_mesh_download(dev, HX_FIELD_RCACHE, c, rcache[c].data(), rcache[c].size()), "download");
    check(hx_mesh_download(dev, HX_FIELD_TSS, c, tss[c].data(), tss[c].size()), "download");
    if (!av[c].empty()) check(hx_mesh_download(dev, HX_FIELD_AV, c, av[c].data(), av[c].size()), "download");
    if (!forcing[c].empty())
      check(hx_mesh_download(dev, HX_FIELD_FORCING, c, forcing[c].data(), forcing[c].size()), "download");
    if (!advection[c].empty())
      check(hx_mesh_download(dev, HX_FIELD_ADVECTION, c, advection[c].data(), advection[c].size()), "download");
  }
  check(hx_mesh_download(dev, HX_FIELD_FACES, 0, faces.data(), faces.size()), "download");
  scatter();
}

// ---- leases -----------------------------------------------------------------------------------------------------

Mesh_lease::Mesh_lease(Sequence<Kernel_element&>& s, int n_dim, int row_size, bool deformed)
{
  am = Accelerated_mesh::find(&s, &c);
  if (am && c >= 0) return;
  // unregistered (or the mixed list): a temporary mesh over these elements alone
  std::vector<Kernel_element*> none, els = collect(s);
  temp.reset(new Accelerated_mesh(n_dim, row_size, deformed ? none : els, deformed ? els : none, {}, {}, {}, false, false));
  am = temp.get();
  c = deformed ? HX_DEFORMED : HX_CARTESIAN;
  am->push();
}

Mesh_lease::Mesh_lease(Sequence<Kernel_connection&>& s, int n_dim, int row_size, bool deformed)
{
  am = Accelerated_mesh::find(&s, &c);
  if (am) return;
  std::vector<Kernel_connection*> none, cs = collect(s);
  temp.reset(new Accelerated_mesh(n_dim, row_size, {}, {}, deformed ? none : cs, deformed ? cs : none, {}, false, false));
  am = temp.get();
  c = deformed ? HX_DEFORMED : HX_CARTESIAN;
  am->push();
}

Mesh_lease::Mesh_lease(Sequence<Refined_face&>& s, int n_dim, int row_size)
{
  am = Accelerated_mesh::find(&s, &c);
  if (am) return;
  temp.reset(new Accelerated_mesh(n_dim, row_size, {}, {}, {}, {}, collect(s), false, false));
  am = temp.get();
  am->push();
}

Mesh_lease::Mesh_lease(Kernel_mesh& mesh)
{
  am = Accelerated_mesh::find(&mesh.car_elems, &c);
  if (am) return;
  temp.reset(new Accelerated_mesh(mesh.n_dim, mesh.row_size, collect(mesh.car_elems), collect(mesh.def_elems),
                                  collect(mesh.car_cons), collect(mesh.def_cons), collect(mesh.ref_faces), true, false));
  am = temp.get();
  am->push();
}

Mesh_lease::~Mesh_lease()
{
  if (temp) {
    try {
      temp->pull();
    } catch (...) {
    }
  }
}

namespace
{
// the face-1 index matching face-0 index i (Spatial.hpp:85-129: transpose, then flip)
int face1_index(int nd, int rs, Connection_direction dir, int i)
{
  if (nd == 3) {
    int c = i / rs, r = i % rs;
    const bool tr = dir.transpose();
    if (dir.flip_tangential()) {
      const bool colwise = (dir.i_dim[0] > 3 - dir.i_dim[0] - dir.i_dim[1]) != tr;
      if (colwise) r = rs - 1 - r;
      else c = rs - 1 - c;
    }
    return tr ? r * rs + c : c * rs + r;
  }
  if (nd == 2) return dir.flip_tangential() ? rs - 1 - i : i;
  return i;
}

class Face_perm : public Face_permutation_dynamic
{
  int nd, rs, nfq, n_var;
  Connection_direction dir;
  double* data;

  public:
  Face_perm(int n_dim, int row_size, Connection_direction d, double* target, int nv)
  : nd{n_dim}, rs{row_size}, nfq{1}, n_var{nv}, dir{d}, data{target}
  {
    for (int i = 1; i < nd; ++i) nfq *= rs;
  }
  void match_faces() override
  {
    std::vector<double> old(data, data + n_var * nfq);
    for (int v = 0; v < n_var; ++v)
      for (int i = 0; i < nfq; ++i) data[v * nfq + i] = old[v * nfq + face1_index(nd, rs, dir, i)];
  }
  void restore() override
  {
    std::vector<double> cur(data, data + n_var * nfq);
    for (int v = 0; v < n_var; ++v)
      for (int i = 0; i < nfq; ++i) data[v * nfq + face1_index(nd, rs, dir, i)] = cur[v * nfq + i];
  }
};
}  // namespace

std::unique_ptr<Face_permutation_dynamic> face_permutation(int n_dim, int row_size, Connection_direction dir,
                                                           double* data, int n_var)
{
  if (n_dim < 1 || n_dim > HX_MAX_N_DIM || row_size < 2 || row_size > HX_MAX_ROW_SIZE)
    throw std::runtime_error("demand for invalid kernel");
  return std::unique_ptr<Face_permutation_dynamic>(new Face_perm(n_dim, row_size, dir, data, n_var));
}

void update(Accelerated_mesh& mesh, const hx_update_options& opts, hx_update_status& status)
{
  check(hx_update(mesh.device(), &opts, &status), "Solver::update");
}

bool is_admissible(Accelerated_mesh& mesh)
{
  int admissible = 0;
  check(hx_is_admissible(mesh.device(), &admissible), "Solver::is_admissible");
  return admissible;
}

int fix_admissibility(Accelerated_mesh& mesh, const hx_update_options& opts, double stability_ratio)
{
  int iters = 0;
  check(hx_fix_admissibility(mesh.device(), &opts, stability_ratio, &iters), "Solver::fix_admissibility");
  return iters;
}

void calc_jacobian(Accelerated_mesh& mesh, const hx_vertex_data& vertices)
{
  check(hx_calc_jacobian(mesh.device(), &vertices), "Solver::calc_jacobian");
}

namespace
{
hx_bc bc_of(int kind)
{
  hx_bc bc{};
  bc.kind = kind;
  bc.coercion = 2.;
  return bc;
}

hx_bc bc_with_state(int kind, const std::vector<double>& state)
{
  if (state.empty() || state.size() > HX_MAX_N_DIM + 2)
    throw std::runtime_error("hexed_amd: freestream state must have n_dim + 2 entries");
  hx_bc bc = bc_of(kind);
  std::copy(state.begin(), state.end(), bc.state);
  return bc;
}
}  // namespace

hx_bc freestream(const std::vector<double>& state) { return bc_with_state(HX_BC_FREESTREAM, state); }
hx_bc riemann_invariants(const std::vector<double>& state) { return bc_with_state(HX_BC_RIEMANN_INVARIANTS, state); }
hx_bc nonpenetration() { return bc_of(HX_BC_NONPENETRATION); }
hx_bc copy() { return bc_of(HX_BC_COPY); }
hx_bc outflow() { return bc_of(HX_BC_OUTFLOW); }

hx_bc pressure_outflow(double pressure)
{
  hx_bc bc = bc_of(HX_BC_PRESSURE_OUTFLOW);
  bc.pressure = pressure;
  return bc;
}

hx_bc no_slip_heat_flux(double heat_flux, double coercion)
{
  hx_bc bc = bc_of(HX_BC_NO_SLIP);
  bc.thermal = HX_THERMAL_HEAT_FLUX;
  bc.thermal_value = heat_flux;
  bc.coercion = coercion;
  return bc;
}

hx_bc no_slip_energy(double energy_per_mass, double coercion)
{
  hx_bc bc = no_slip_heat_flux(0., coercion);
  bc.thermal = HX_THERMAL_ENERGY;
  bc.thermal_value = energy_per_mass;
  return bc;
}

hx_bc no_slip_equilibrium(double emissivity, double heat_transfer_coef, double temperature, double coercion)
{
  hx_bc bc = no_slip_heat_flux(0., coercion);
  bc.thermal = HX_THERMAL_EQUILIBRIUM;
  bc.emissivity = emissivity;
  bc.heat_transfer_coef = heat_transfer_coef;
  bc.temperature = temperature;
  return bc;
}

}  // namespace hexed::amd

namespace hexed
{

namespace
{

hx_kernel_options abi_opts(const Kernel_options& o) { return {o.dt, o.i_stage, o.compute_residual, o.use_filter}; }

void count(Kernel_mesh& mesh, const Kernel_options& opts, bool neighbor)
{
  // Stopwatch_tree work units of the reference's timed kernel calls (kernel_factory.hpp:32-45)
  auto add = [](Stopwatch_tree& t, const char* name, long long n) {
    auto it = t.children.find(name);
    if (it != t.children.end()) it->second.work_units_completed += n;
  };
  if (neighbor) {
    add(opts.sw_car, "neighbor", mesh.car_cons.size());
    add(opts.sw_def, "neighbor", mesh.def_cons.size());
  }
  add(opts.sw_car, "local", mesh.car_elems.size());
  add(opts.sw_def, "local", mesh.def_elems.size());
  opts.sw_pr.work_units_completed += mesh.ref_faces.size();
}

}  // namespace

// compute_euler (kernels_convective.cpp:8-18)
void compute_euler(Kernel_mesh mesh, Kernel_options opts)
{
  if (opts.i_stage && opts.compute_residual) throw std::runtime_error("residual calculation is a single-stage operation");
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  amd::check(hx_compute_euler(m.device(), &o), "compute_euler");
  count(mesh, opts, true);
}

namespace
{
// COMPUTE_DIFFUSION (kernels_diffusive.cpp:8-30) with a host flux_bc callback, which needs the face data on the host:
// the device sequence is split around it
void diffusion_with_host_bc(Kernel_mesh& mesh, amd::Mesh_lease& m, const hx_kernel_options& o, const hx_pde& p,
                            const std::function<void()>& flux_bc)
{
  hx_mesh* d = m.device();
  amd::check(hx_neighbor(d, HX_CARTESIAN, &p, 0), "neighbor");
  amd::check(hx_neighbor(d, HX_DEFORMED, &p, 0), "neighbor");
  amd::check(hx_restrict_refined(d, &p, 1, 0), "restrict");
  amd::check(hx_restrict_refined(d, &p, 0, 1), "restrict");
  amd::check(hx_local(d, HX_CARTESIAN, &p, o.dt, 0, o.compute_residual, o.use_filter), "local");
  amd::check(hx_local(d, HX_DEFORMED, &p, o.dt, 0, o.compute_residual, o.use_filter), "local");
  amd::check(hx_prolong_refined(d, &p, 1, 1), "prolong");
  int cls;
  amd::Accelerated_mesh* am = amd::Accelerated_mesh::find(&mesh.car_elems, &cls);
  if (am) am->pull();
  flux_bc();
  if (am) am->push();
  amd::check(hx_neighbor_reconcile(d, HX_CARTESIAN, &p), "neighbor_reconcile");
  amd::check(hx_neighbor_reconcile(d, HX_DEFORMED, &p), "neighbor_reconcile");
  amd::check(hx_restrict_refined(d, &p, 1, 1), "restrict");
  amd::check(hx_reconcile_ldg_flux(d, HX_CARTESIAN, &p, o.dt, 0, o.compute_residual, o.use_filter), "reconcile");
  amd::check(hx_reconcile_ldg_flux(d, HX_DEFORMED, &p, o.dt, 0, o.compute_residual, o.use_filter), "reconcile");
  amd::check(hx_prolong_refined(d, &p, 0, 0), "prolong");
}

}  // namespace

// compute_navier_stokes (kernels_diffusive.cpp:8-29). A host flux_bc callback needs the face data on the host, so
// with one the device sequence is split around it; without one, the mesh's device boundary conditions apply.
void compute_navier_stokes(Kernel_mesh mesh, Kernel_options opts, std::function<void()> flux_bc, Transport_model visc,
                           Transport_model therm_cond)
{
  if (opts.i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  if (!flux_bc) {
    amd::check(hx_compute_navier_stokes(m.device(), &o, visc.abi(), therm_cond.abi()), "compute_navier_stokes");
  } else {
    diffusion_with_host_bc(mesh, m, o, hx_pde{HX_PDE_NAVIER_STOKES, visc.abi(), therm_cond.abi(), 0., 0., 0.}, flux_bc);
  }
  count(mesh, opts, true);
}

// compute_advection (kernels_convective.cpp:8-19 with pde::Advection)
void compute_advection(Kernel_mesh mesh, Kernel_options opts, double advect_length)
{
  if (opts.i_stage && opts.compute_residual) throw std::runtime_error("residual calculation is a single-stage operation");
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  amd::check(hx_compute_advection(m.device(), &o, advect_length), "compute_advection");
  count(mesh, opts, true);
}

// compute_smooth_av (kernels_diffusive.cpp:31)
void compute_smooth_av(Kernel_mesh mesh, Kernel_options opts, std::function<void()> flux_bc, double diff_time,
                       double chebyshev_step)
{
  if (opts.i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  if (!flux_bc) amd::check(hx_compute_smooth_av(m.device(), &o, diff_time, chebyshev_step), "compute_smooth_av");
  else diffusion_with_host_bc(mesh, m, o, hx_pde{HX_PDE_SMOOTH_AV, inviscid.abi(), inviscid.abi(), 0., diff_time, chebyshev_step},
                              flux_bc);
  count(mesh, opts, true);
}

// compute_fix_therm_admis (kernels_diffusive.cpp:32)
void compute_fix_therm_admis(Kernel_mesh mesh, Kernel_options opts, std::function<void()> flux_bc)
{
  if (opts.i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  if (!flux_bc) amd::check(hx_compute_fix_therm_admis(m.device(), &o), "compute_fix_therm_admis");
  else diffusion_with_host_bc(mesh, m, o, hx_pde{HX_PDE_FIX_ADMIS, inviscid.abi(), inviscid.abi(), 0., 0., 0.}, flux_bc);
  count(mesh, opts, true);
}

// max_dt_advection / max_dt_smooth_av / max_dt_fix_therm_admis (kernels_max_dt.cpp:13-21)
double max_dt_advection(Kernel_mesh mesh, Kernel_options opts, double msc, double msd, bool local_time,
                        double advect_length)
{
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  double dt = 0.;
  amd::check(hx_max_dt_advection(m.device(), &o, msc, msd, local_time, advect_length, &dt), "max_dt_advection");
  return dt;
}

double max_dt_smooth_av(Kernel_mesh mesh, Kernel_options opts, double msc, double msd, bool local_time)
{
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  double dt = 0.;
  amd::check(hx_max_dt_smooth_av(m.device(), &o, msc, msd, local_time, &dt), "max_dt_smooth_av");
  return dt;
}

double max_dt_fix_therm_admis(Kernel_mesh mesh, Kernel_options opts, double msc, double msd, bool local_time)
{
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  double dt = 0.;
  amd::check(hx_max_dt_fix_therm_admis(m.device(), &o, msc, msd, local_time, &dt), "max_dt_fix_therm_admis");
  return dt;
}

// kernels_convective.cpp:38-56; Write_face<Advection>(basis, 1.) and Write_face<Smooth_art_visc>(basis, 1., 1.)
void compute_prolong_advection(Kernel_mesh mesh)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_prolong_advection(m.device(), 1.), "compute_prolong_advection");
}

void compute_write_face_advection(Kernel_mesh mesh)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_write_face_advection(m.device(), 1.), "compute_write_face_advection");
}

void compute_write_face_smooth_av(Kernel_mesh mesh)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_write_face_smooth_av(m.device()), "compute_write_face_smooth_av");
}

std::unique_ptr<Face_permutation_dynamic> face_permutation(int n_dim, int row_size, Connection_direction dir, double* data)
{
  return amd::face_permutation(n_dim, row_size, dir, data, n_dim + 2);  // Navier_stokes<false>'s n_extrap
}

// max_dt_euler / max_dt_navier_stokes (kernels_max_dt.cpp:8-17)
double max_dt_euler(Kernel_mesh mesh, Kernel_options opts, double msc, double msd, bool local_time)
{
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  double dt = 0.;
  amd::check(hx_max_dt_euler(m.device(), &o, msc, msd, local_time, &dt), "max_dt_euler");
  return dt;
}

double max_dt_navier_stokes(Kernel_mesh mesh, Kernel_options opts, double msc, double msd, bool local_time,
                            Transport_model visc, Transport_model therm_cond)
{
  amd::Mesh_lease m(mesh);
  hx_kernel_options o = abi_opts(opts);
  double dt = 0.;
  amd::check(hx_max_dt_navier_stokes(m.device(), &o, msc, msd, local_time, visc.abi(), therm_cond.abi(), &dt),
             "max_dt_navier_stokes");
  return dt;
}

void compute_prolong(Kernel_mesh mesh, bool scale, bool offset)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_prolong(m.device(), scale, offset), "compute_prolong");
}

void compute_restrict(Kernel_mesh mesh, bool scale, bool offset)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_restrict(m.device(), scale, offset), "compute_restrict");
}

void compute_write_face(Kernel_mesh mesh)
{
  amd::Mesh_lease m(mesh);
  amd::check(hx_compute_write_face(m.device()), "compute_write_face");
}

}  // namespace hexed
