// C ABI of include/hexed_amd.h: contexts, device meshes, and the kernel sequences of the reference's free
// functions (kernels_convective.cpp, kernels_diffusive.cpp, kernels_max_dt.cpp) and Solver::update.
#include <algorithm>
#include <cfloat>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "basis_tables.inc"
#include "hx_launch.hpp"

using namespace hx;

// Scratch memory of hx_update_no_alloc, for callers that cannot allocate device memory: [0, 32) holds the reduction
// scalar of the time step, [32, 32 + no_alloc_max_blocks) one fused Max_dt minimum per Euler Local workgroup (a mesh
// of up to 262,144 workgroups: 262,144 elements at 3D row size 5 to 8). Each GPU holds its own copy (2 MiB) until the
// process ends. Host code uses the address hipGetSymbolAddress returns for the current device (hx_ctx::scratch), never
// the array's name.
namespace {
constexpr long long no_alloc_max_blocks = 1LL << 18;
constexpr long long scratch_minima_offset = 32;
}  // namespace
static __device__ __align__(256) double scratch[scratch_minima_offset + no_alloc_max_blocks];

namespace {
constexpr int max_devices = 64;
std::mutex scratch_mutex;
double* scratch_address[max_devices] = {};       // per device, once found
hipError_t scratch_error[max_devices] = {};
// advanced by every hx_update_no_alloc on the device: a mesh's fused minima in the array are current while the
// epoch is the one its own last call left
std::atomic<long long> scratch_epochs[max_devices];

// the scratch array's address on `device` (the current device), looked up once per device
double* scratch_on(int device, hipError_t* err)
{
  if (device < 0 || device >= max_devices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device] && scratch_error[device] == hipSuccess) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
      scratch_address[device] = static_cast<double*>(p);
    } else {
      scratch_error[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
      (void)hipGetLastError();
    }
  }
  *err = scratch_error[device];
  return scratch_address[device];
}

long long scratch_epoch(int device) { return scratch_epochs[device].load(); }
}  // namespace

struct hx_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* ops[HX_MAX_ROW_SIZE + 1] = {};  // device Ops<rs> tables
  std::vector<char> host_ops[HX_MAX_ROW_SIZE + 1];  // host copies, for kernels taking constants by value
  void* eops[HX_MAX_ROW_SIZE + 1] = {};             // device Eops<rs, double> tables
  double* scalar = nullptr;             // device scratch for reductions
  double* host_scalar = nullptr;        // pinned host memory the per-step reductions are read back into
  // per-category event timing
  struct Pending
  {
    hipEvent_t start, stop;
    std::string category;
    long long units;
  };
  struct Total
  {
    double ms = 0;
    long long launches = 0, units = 0;
  };
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<Pending> pending;
  std::map<std::string, Total> totals;
  std::vector<hx_mesh*> meshes;  // live meshes of this context, detached by hx_ctx_destroy
  std::vector<hx_group*> groups; // live groups of this context, detached by hx_ctx_destroy
  bool events_off = false;       // no per-launch events (a graph capture, or a call that may not create any)
  double* scratch = nullptr;     // this device's address of the file-scope scratch array (hx_update_no_alloc)
  hipError_t scratch_err = hipSuccess;
};

// status of the updates enqueued by hx_update_async, kept on the device
struct Step_dev
{
  double dt;          // the time step the stage kernels of the current step read
  double time_step;   // hx_update_status::time_step
  double flow_time;   // hx_update_status::flow_time
  long long iteration;
};

struct hx_mesh
{
  hx_ctx* ctx = nullptr;
  Dev_mesh dm{};
  std::vector<void*> allocs;
  const Kernel_set* ks = nullptr;
  int precision = HX_FP64;
  // batch of independent cases (hx_mesh_set_cases): case of every element per class, and per case the Max_dt
  // minimum (bits of a positive double) and the time step
  int n_cases = 0;
  int* case_id[2] = {};
  unsigned long long* case_min = nullptr;
  double* case_dt = nullptr;
  // when each class lists its cases in order (non-decreasing ids): first element of every case, [n_cases + 1]
  int* case_start[2] = {};
  int case_split[2] = {};  // workgroups per case of the per-case reduction
  // Max_dt fused into the last stage of hx_update: every call that can change device data (a kernel launch, an
  // upload, handing out a device pointer) advances `gen`; the next hx_update takes the fused minimum only if `gen` is
  // where the previous hx_update left it and its time-step settings are the same
  long long gen = 0;
  double* dt_block_min = nullptr;  // one minimum per workgroup of the Euler Local, both classes
  long long n_dt_blocks = 0;
  Dt_fuse fuse{};                  // the fusion the next Euler Local launch performs (block_min == nullptr: none)
  int* admis_flags = nullptr;      // [2] results of the admissibility check (hx_is_admissible)
  struct
  {
    bool valid = false;
    long long gen = 0;
    double msc = 0, msd = 0;
    int use_ldg = 0;
    hx_transport visc{}, cond{};
    bool in_scratch = false;  // the minima are in the scratch array (hx_update_no_alloc), not in dt_block_min
  } dt_cache;
  // the same fusion for hx_update_cases (cases listed in order): stage 1's Local writes each point's 1/scale into
  // case_scale, which the next call's per-case minimum reads instead of running Max_dt
  double* case_scale[2] = {};
  struct
  {
    bool valid = false;
    long long gen = 0;
    double cc = 0, cd = 0;
    int use_ldg = 0;
    hx_transport visc{}, cond{};
  } case_cache;
  // hx_update_async: the device status, the time step the stage kernels read while an update is enqueued with it
  // (null otherwise), and the captured steady-state update with the options and form it was captured for
  Step_dev* step = nullptr;
  const double* dt_dev = nullptr;
  // the reduction scalar the mesh's calls use: the context's, or the file-scope scratch array during
  // hx_update_no_alloc; and the scratch epoch of the last hx_update_no_alloc on this mesh (its fused minima sit in the
  // scratch array and are current only while no other call has used the array since)
  double* scalar = nullptr;
  long long scratch_epoch = -1;
  int n_groups = 0;  // groups (hx_group) holding the mesh, which must be destroyed first
  // face-flux folding of the current stage of an update (Dev_mesh::partner): the buffer holding the face states the
  // stage reads (alt_rd) and the one its face extrapolations go to (alt_wr)
  int fold = 0, alt_rd = 0, alt_wr = 0;
  int face_fold_mode = -1;  // hx_mesh_set_face_fold: 1 fold, 0 do not, -1 fold where rs >= 2 n_dim
  bool mortar_fusable = false;  // every refined face's mortars are in connections (k_mortar_restrict applies)
  hipGraphExec_t graph = nullptr;
  hx_update_options graph_opts{};
  int graph_precision = -1;
  // the vertex topology of the last call that shared vertex values (hx_calc_jacobian, hx_set_art_visc_admis): a host
  // copy, against which a repeated topology is recognised and neither validated nor uploaded again, its device copy,
  // and the device work arrays of the vertex sharing and of the element values, kept for the next call
  struct
  {
    std::vector<int> ids;      // vertex ids of the Cartesian elements, then of the deformed ones
    std::vector<int> hanging;  // hx_vertex_data::hanging
    long long n_vertex = -1;   // -1: nothing cached
    int* vid = nullptr;
    int* hang = nullptr;
    double* shared = nullptr;  // [n_vertex]
    double* vals = nullptr;    // [9 n_hanging]
    double* vert = nullptr;    // [n_cartesian + n_deformed][2^n_dim] per-element vertex copies
    double* elem = nullptr;    // [n_cartesian + n_deformed] element values
    long long cap_vertex = 0, cap_hanging = 0;  // sizes `shared` and `hang` / `vals` are allocated for
  } vc;
};


// Meshes stepped together (hx_group_update): each kernel of an Euler stage launched once for all of them, both classes
// and every row size, when every mesh is 3D with row size 4, 5 or 6 (the fused set); otherwise the meshes step one
// after the other.
struct hx_group
{
  hx_ctx* ctx = nullptr;
  std::vector<hx_mesh*> meshes;
  const hx::Fused_set* fs = nullptr;    // null: no fused launches
  bool with_def = false;                // some mesh has deformed elements
  int n_cases = 0;                      // over all meshes
  hx::Fused_member* mem = nullptr;      // device [2][n]: stage 0 (no fusion), stage 1 (the batch's fused time step)
  hx::Fused_plan* plans = nullptr;      // device [3]: state BCs, Neighbor, Local
  hx::Fused_plan hplan[3]{};
  std::vector<hx::Fused_member> hmem;   // host copy of `mem`
  double* dts = nullptr;                // pinned host: every case's time step
  int order = 2, per_cu = 0;            // launch shape of the fused Local (hx_group_set_shape)
  int n_cu = 0;
  const void* hops[HX_MAX_ROW_SIZE + 1] = {};  // the context's host Ops<rs> tables, by row size
};

namespace {

// a mesh that can run work: not NULL and its context not destroyed
bool live(const hx_mesh* mesh) { return mesh && mesh->ctx; }

// frees a mesh's device memory; the struct stays valid and holds no device pointer afterwards
void release_graph(hx_mesh* mesh)
{
  if (mesh->graph) (void)hipGraphExecDestroy(mesh->graph);
  mesh->graph = nullptr;
  mesh->graph_precision = -1;
}

void release_vertex_cache(hx_mesh* mesh)
{
  auto& vc = mesh->vc;
  for (void* p : {(void*)vc.vid, (void*)vc.hang, (void*)vc.shared, (void*)vc.vals, (void*)vc.vert, (void*)vc.elem})
    if (p) (void)hipFree(p);
  vc.vid = vc.hang = nullptr;
  vc.shared = vc.vals = vc.vert = vc.elem = nullptr;
  vc.ids.clear();
  vc.hanging.clear();
  vc.n_vertex = -1;
  vc.cap_vertex = vc.cap_hanging = 0;
}

void release_mesh_memory(hx_mesh* mesh)
{
  release_graph(mesh);
  release_vertex_cache(mesh);
  for (void* p : mesh->allocs) (void)hipFree(p);
  mesh->allocs.clear();
  mesh->dm = Dev_mesh{};
  mesh->n_cases = 0;
  mesh->case_id[0] = mesh->case_id[1] = nullptr;
  mesh->case_start[0] = mesh->case_start[1] = nullptr;
  mesh->case_scale[0] = mesh->case_scale[1] = nullptr;
  mesh->case_min = nullptr;
  mesh->case_dt = nullptr;
  mesh->dt_block_min = nullptr;
  mesh->admis_flags = nullptr;
  mesh->step = nullptr;
  mesh->dt_dev = nullptr;
  mesh->fuse = Dt_fuse{};
  mesh->dt_cache.valid = mesh->case_cache.valid = false;
}

// frees a group's memory and releases its meshes; the struct stays valid for hx_group_destroy
void release_group(hx_group* g)
{
  if (g->mem) (void)hipFree(g->mem);
  if (g->plans) (void)hipFree(g->plans);
  if (g->dts) (void)hipHostFree(g->dts);
  g->mem = nullptr;
  g->plans = nullptr;
  g->dts = nullptr;
  for (hx_mesh* m : g->meshes) --m->n_groups;
  g->meshes.clear();
  g->ctx = nullptr;
}

hx_status fail(hx_ctx* ctx, hx_status code, const std::string& msg)
{
  if (ctx) ctx->err = msg;
  return code;
}

hx_status hip_check(hx_ctx* ctx, hipError_t e, const char* what)
{
  if (e == hipSuccess) return HX_OK;
  return fail(ctx, HX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HX_TRY(ctx, expr, what)                                     \
  do {                                                              \
    hx_status s_ = hip_check((ctx), (expr), (what));                \
    if (s_ != HX_OK) return s_;                                     \
  } while (0)

template <int RS>
void fill_ops(Ops<RS>& o, int rs_index)
{
  using namespace hx_basis_tables;
  const double* d = diff_mats[rs_index];
  const double* b = boundarys[rs_index];
  const double* w = weights[rs_index];
  const double* f = filters[rs_index];
  const double* p = prolongs[rs_index];
  const double* r = restricts[rs_index];
  for (int i = 0; i < RS; ++i) {
    for (int j = 0; j < RS; ++j) {
      o.diff[i][j] = d[i * RS + j];
      o.filt[i][j] = f[i * RS + j];
      for (int h = 0; h < 2; ++h) {
        o.prolong[h][i][j] = p[(h * RS + i) * RS + j];
        o.restrict_[h][i][j] = r[(h * RS + i) * RS + j];
      }
    }
    for (int s = 0; s < 2; ++s) o.bound[s][i] = b[s * RS + i];
    // Derivative.hpp:27-28: lift = diag(1/w) * boundary^T * diag(-1, 1)
    const double inv_w = 1. / w[i];
    o.lift[i][0] = -(inv_w * b[i]);
    o.lift[i][1] = inv_w * b[RS + i];
    o.node[i] = nodes[rs_index][i];
    o.weight[i] = w[i];
    for (int k = 0; k < RS; ++k) o.orth[k][i] = orthogonals[rs_index][k * RS + i];
  }
}

template <int RS>
hx_status upload_ops(hx_ctx* ctx, void** out)
{
  Ops<RS> o;
  fill_ops<RS>(o, RS - 2);
  HX_TRY(ctx, hipMalloc(out, sizeof(o)), "hipMalloc(ops)");
  HX_TRY(ctx, hipMemcpy(*out, &o, sizeof(o), hipMemcpyHostToDevice), "hipMemcpy(ops)");
  const char* bytes = reinterpret_cast<const char*>(&o);
  ctx->host_ops[RS].assign(bytes, bytes + sizeof(o));
  const Eops<RS, double> e = make_eops<RS, double>(o);
  HX_TRY(ctx, hipMalloc(&ctx->eops[RS], sizeof(e)), "hipMalloc(eops)");
  HX_TRY(ctx, hipMemcpy(ctx->eops[RS], &e, sizeof(e), hipMemcpyHostToDevice), "hipMemcpy(eops)");
  return HX_OK;
}

hx_status ensure_ops(hx_ctx* ctx, int rs)
{
  if (ctx->ops[rs]) return HX_OK;
  switch (rs) {
    case 2: return upload_ops<2>(ctx, &ctx->ops[2]);
    case 3: return upload_ops<3>(ctx, &ctx->ops[3]);
    case 4: return upload_ops<4>(ctx, &ctx->ops[4]);
    case 5: return upload_ops<5>(ctx, &ctx->ops[5]);
    case 6: return upload_ops<6>(ctx, &ctx->ops[6]);
    case 7: return upload_ops<7>(ctx, &ctx->ops[7]);
    case 8: return upload_ops<8>(ctx, &ctx->ops[8]);
  }
  return fail(ctx, HX_ERR_INVALID, "demand for invalid kernel");
}

const Kernel_set* kernel_set(int nd, int rs)
{
  switch (nd) {
    case 1: return kernel_set_nd1(rs);
    case 2: return kernel_set_nd2(rs);
    case 3: return kernel_set_nd3(rs);
  }
  return nullptr;
}

int ipow_h(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

struct Sizes
{
  int nd, rs, nq, nfq, nv, nvert, face_sz;
  explicit Sizes(int n_dim, int row_size)
  : nd{n_dim}, rs{row_size}, nq{ipow_h(row_size, n_dim)}, nfq{ipow_h(row_size, n_dim - 1)}, nv{n_dim + 2},
    nvert{1 << n_dim}, face_sz{std::max(2 * (n_dim + 2), n_dim + row_size) * ipow_h(row_size, n_dim - 1)}
  {}
};

// number of doubles of a field of class cls, or -1 for an unknown field
long long field_size(const Dev_mesh& m, int field, int cls)
{
  const long long n = (cls == 0 || cls == 1) ? m.n_elem[cls] : 0;
  switch (field) {
    case HX_FIELD_STATE: return n * m.nv * m.nq;
    case HX_FIELD_RCACHE: return n * std::max(m.nv, m.rs) * m.nq;
    case HX_FIELD_TSS: return n * m.nq;
    case HX_FIELD_AV: return n * 2 * m.nq;
    case HX_FIELD_FORCING: return n * 4 * m.nq;
    case HX_FIELD_ADVECTION: return n * m.rs * m.nq;
    case HX_FIELD_VERTEX_TSS: return n * (1 << m.nd);
    case HX_FIELD_METRICS: return cls == HX_DEFORMED ? n * (m.nd * m.nd + 1) * m.nq : -1;
    case HX_FIELD_FACES: return (long long)m.n_slot * m.face_sz;
    case HX_FIELD_NORMALS: return (long long)m.n_nslot * m.nd * m.nfq;
  }
  return -1;
}

double** field_ptr(Dev_mesh& m, int field, int cls)
{
  const bool c_ok = cls == 0 || cls == 1;
  switch (field) {
    case HX_FIELD_STATE: return c_ok ? &m.state[cls] : nullptr;
    case HX_FIELD_RCACHE: return c_ok ? &m.rcache[cls] : nullptr;
    case HX_FIELD_TSS: return c_ok ? &m.tss[cls] : nullptr;
    case HX_FIELD_AV: return c_ok ? &m.av[cls] : nullptr;
    case HX_FIELD_FORCING: return c_ok ? &m.forcing[cls] : nullptr;
    case HX_FIELD_ADVECTION: return c_ok ? &m.advection[cls] : nullptr;
    case HX_FIELD_VERTEX_TSS: return c_ok ? &m.vertex_tss[cls] : nullptr;
    case HX_FIELD_METRICS: return cls == HX_DEFORMED ? &m.metrics : nullptr;
    case HX_FIELD_FACES: return &m.faces;
    case HX_FIELD_NORMALS: return &m.normals;
  }
  return nullptr;
}

template <typename T>
hx_status device_copy(hx_mesh* mesh, T** dst, const T* src, long long count, bool zero_if_null)
{
  *dst = nullptr;
  if (count <= 0) return HX_OK;
  if (!src && !zero_if_null) return HX_OK;
  void* p = nullptr;
  HX_TRY(mesh->ctx, hipMalloc(&p, size_t(count) * sizeof(T)), "hipMalloc(mesh)");
  mesh->allocs.push_back(p);
  if (src) HX_TRY(mesh->ctx, hipMemcpy(p, src, size_t(count) * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(mesh)");
  else HX_TRY(mesh->ctx, hipMemset(p, 0, size_t(count) * sizeof(T)), "hipMemset(mesh)");
  *dst = static_cast<T*>(p);
  return HX_OK;
}

// Every index a kernel dereferences is checked here, so a malformed mesh is an error rather than a fault.
std::string validate(const hx_host_mesh* d, const Sizes& z)
{
  auto in = [](long long v, long long lo, long long hi) { return v >= lo && v < hi; };
  for (int c = 0; c < 2; ++c) {
    if (d->n_elem[c] < 0 || d->n_con[c] < 0) return "negative element or connection count";
    if (d->n_elem[c] && (!d->state[c] || !d->nom_size[c] || !d->vertex_tss[c] || !d->face_slot[c]))
      return "element arrays missing";
    for (long long k = 0; k < (long long)d->n_elem[c] * 2 * z.nd; ++k)
      if (!in(d->face_slot[c][k], 0, d->n_slot)) return "face slot out of range";
    if (d->n_con[c] && (!d->con_slot[c] || !d->con_dir[c])) return "connection arrays missing";
    for (long long k = 0; k < (long long)d->n_con[c]; ++k) {
      if (!in(d->con_slot[c][2 * k], 0, d->n_slot) || !in(d->con_slot[c][2 * k + 1], 0, d->n_slot))
        return "connection slot out of range";
      int p = d->con_dir[c][k];
      if ((p & 3) >= z.nd || ((p >> 2) & 3) >= z.nd || (p >> 6) != 0) return "connection direction out of range";
    }
  }
  if (d->n_elem[HX_DEFORMED]) {
    if (!d->metrics || !d->face_nslot) return "deformed metrics missing";
    for (long long k = 0; k < (long long)d->n_elem[1] * 2 * z.nd; ++k)
      if (!in(d->face_nslot[k], -1, d->n_nslot)) return "face normal slot out of range";
  }
  if (d->n_con[HX_DEFORMED] && (!d->con_nslot || !d->normals)) return "deformed connection normals missing";
  if (d->n_slot && !d->faces) return "face buffer missing";
  if (d->n_ref && (!d->ref_slot || !d->ref_stretch)) return "refined face arrays missing";
  if (d->n_bc_con && (!d->bc_con || !d->bc_serial || !d->bcs)) return "boundary arrays missing";
  for (long long k = 0; k < 2LL * d->n_con[HX_DEFORMED]; ++k)
    if (!in(d->con_nslot[k], 0, d->n_nslot)) return "connection normal slot out of range";
  for (long long k = 0; k < d->n_ref; ++k) {
    if (!in(d->ref_slot[5 * k], 0, d->n_slot)) return "refined coarse slot out of range";
    int nf = 1 << (z.nd - 1);
    for (int dd = 0; dd < z.nd - 1; ++dd) nf /= 1 + (d->ref_stretch[2 * k + dd] != 0);
    for (int f = 0; f < nf; ++f)
      if (!in(d->ref_slot[5 * k + 1 + f], 0, d->n_slot)) return "refined fine slot out of range";
  }
  for (long long k = 0; k < d->n_bc_con; ++k) {
    if (!in(d->bc_con[k], 0, d->n_con[HX_DEFORMED])) return "boundary connection out of range";
    if (!in(d->bc_serial[k], 0, d->n_bc)) return "boundary condition serial out of range";
  }
  for (long long k = 0; k < d->n_bc; ++k) {
    if (!in(d->bcs[k].kind, HX_BC_FREESTREAM, HX_BC_NO_SLIP + 1)) return "unknown boundary condition kind";
    if (d->bcs[k].kind == HX_BC_NO_SLIP && !in(d->bcs[k].thermal, HX_THERMAL_HEAT_FLUX, HX_THERMAL_EQUILIBRIUM + 1))
      return "unknown thermal condition";
  }
  if (d->n_slot < 0 || d->n_nslot < 0 || d->n_ref < 0 || d->n_bc_con < 0 || d->n_bc < 0) return "negative count";
  return "";
}

// The face-flux folding tables (Dev_mesh::partner, mortar_con) of a validated mesh. Every element face in a connection
// that touches no hanging-face mortar gets its neighbour slot, the packed direction with its side, the connection class
// and a boundary bit, and (deformed connections) the connection's normal slot; every other face (mortar neighbours,
// coarse hanging faces) is marked precomputed. Connections touching a mortar are listed per class.
// ref_con[4 * i + f]: the connection (class << 30 | index) of mortar f of refined face i, -1 where the face has none.
// Returns whether the mortars of every refined face are exactly the fine faces its stretch leaves (Restrict_refined's)
// and each is in a connection, which k_mortar_restrict requires.
bool build_partners(const hx_host_mesh* d, const Sizes& z, std::vector<int> (&partner)[2], std::vector<int> (&mortar)[2],
                    std::vector<int>& ref_con)
{
  const int nf = 2 * z.nd;
  std::vector<int> owner(std::max(d->n_slot, 1), -1);  // element face owning each slot: cls | (e * 2nd + face) << 1
  for (int c = 0; c < 2; ++c) {
    partner[c].assign(size_t(d->n_elem[c]) * nf * 4, 0);
    for (long long k = 0; k < (long long)d->n_elem[c] * nf; ++k) {
      partner[c][4 * k] = -1;
      owner[d->face_slot[c][k]] = c | int(k << 1);
    }
  }
  std::vector<char> is_mortar(std::max(d->n_slot, 1), 0), is_bc(std::max(d->n_con[HX_DEFORMED], 1), 0);
  for (int i = 0; i < d->n_ref; ++i)
    for (int f = 1; f < 5; ++f)
      if (d->ref_slot[5 * i + f] >= 0) is_mortar[d->ref_slot[5 * i + f]] = 1;
  for (int i = 0; i < d->n_bc_con; ++i) is_bc[d->bc_con[i]] = 1;
  std::vector<int> con_of_slot(std::max(d->n_slot, 1), -1);  // the connection of each mortar slot
  for (int cc = 0; cc < 2; ++cc) {
    mortar[cc].clear();
    for (int con = 0; con < d->n_con[cc]; ++con) {
      const int s0 = d->con_slot[cc][2 * con], s1 = d->con_slot[cc][2 * con + 1];
      if (is_mortar[s0] || is_mortar[s1]) {
        mortar[cc].push_back(con);
        con_of_slot[is_mortar[s0] ? s0 : s1] = cc << 30 | con;
        continue;
      }
      const int bc = cc == HX_DEFORMED ? is_bc[con] : 0;
      for (int side = 0; side < 2; ++side) {
        const int o = owner[side ? s1 : s0];
        if (o < 0) continue;
        int* p = &partner[o & 1][4 * size_t(o >> 1)];
        p[0] = side ? s0 : s1;
        p[1] = (d->con_dir[cc][con] & 63) | side << 6 | cc << 7 | bc << 8;
        p[2] = cc == HX_DEFORMED ? d->con_nslot[2 * con] : -1;
      }
    }
  }
  ref_con.assign(size_t(d->n_ref) * 4, -1);
  bool complete = true;
  for (int i = 0; i < d->n_ref; ++i) {
    int nf = 1 << (z.nd - 1);
    for (int k = 0; k < z.nd - 1 && k < 2; ++k)
      if (d->ref_stretch[2 * i + k]) nf /= 2;
    for (int f = 0; f < 4; ++f) {
      const int slot = d->ref_slot[5 * i + 1 + f];
      if (slot >= 0) ref_con[4 * size_t(i) + f] = con_of_slot[slot];
      // (a mortar past the stretch's fine faces would not be restricted but still needs its flux)
      complete = complete && (f < nf ? ref_con[4 * size_t(i) + f] >= 0 : slot < 0);
    }
  }
  return complete;
}

// `writes`: the launch may change mesh data (every launch except read-only checks), which retires a fused Max_dt
Launch make_launch(hx_mesh* mesh, bool writes = true)
{
  if (writes) ++mesh->gen;
  Launch l{};
  l.stream = mesh->ctx->stream;
  l.m = mesh->dm;
  l.ops = mesh->ctx->ops[mesh->dm.rs];
  l.hops = mesh->ctx->host_ops[mesh->dm.rs].data();
  l.eops = mesh->ctx->eops[mesh->dm.rs];
  l.scalar = mesh->scalar ? mesh->scalar : mesh->ctx->scalar;
  l.precision = mesh->precision;
  l.fuse = mesh->fuse;
  l.dtp = mesh->dt_dev;
  l.fold = mesh->fold;
  l.m.alt_rd = mesh->alt_rd;
  l.m.alt_wr = mesh->alt_wr;
  return l;
}

Transport to_dev(const hx_transport& t)
{
  return Transport{t.const_val, t.ref_val, t.ref_temp, std::sqrt(t.ref_temp), t.temp_offset};
}

int pde_kind(const hx_pde* pde) { return pde ? pde->kind : HX_PDE_EULER; }
bool is_av(int kind) { return kind == HX_PDE_ADVECTION || kind == HX_PDE_SMOOTH_AV; }
// PDEs served by the generator's kernel family (hx_av_kernels.hpp): the AV generator's two and Fix_therm_admis
bool is_aux(int kind) { return is_av(kind) || kind == HX_PDE_FIX_ADMIS; }

hx_status pde_check(hx_mesh* mesh, const hx_pde* pde, bool* visc, Launch* l)
{
  int kind = pde_kind(pde);
  if (kind != HX_PDE_EULER && kind != HX_PDE_NAVIER_STOKES && !is_aux(kind))
    return fail(mesh->ctx, HX_ERR_UNSUPPORTED, "PDE kind not built into this library");
  if (is_av(kind)) {
    for (int c = 0; c < 2; ++c)
      if (mesh->dm.n_elem[c] && (!mesh->dm.advection[c] || !mesh->dm.forcing[c]))
        return fail(mesh->ctx, HX_ERR_INVALID, "artificial-viscosity PDEs need the forcing and advection fields");
    if (kind == HX_PDE_ADVECTION && !(pde->advect_length > 0))
      return fail(mesh->ctx, HX_ERR_INVALID, "advect_length must be positive");
    if (kind == HX_PDE_SMOOTH_AV && !(pde->diff_time > 0))
      return fail(mesh->ctx, HX_ERR_INVALID, "diff_time must be positive");
  }
  *visc = kind == HX_PDE_NAVIER_STOKES;
  if (*visc) l->ns = Ns_params{to_dev(pde->visc), to_dev(pde->cond)};
  return HX_OK;
}

// Pde::n_extrap: the variables Write_face / Prolong_refined / Restrict_refined move (Spatial.hpp:141,218)
int n_extrap(const hx_mesh* mesh, const hx_pde* pde)
{
  const int kind = pde_kind(pde);
  if (kind == HX_PDE_ADVECTION) return mesh->dm.nd + mesh->dm.rs;
  if (kind == HX_PDE_SMOOTH_AV) return 3;
  return mesh->dm.nv;
}

bool valid_cls(int cls) { return cls == HX_CARTESIAN || cls == HX_DEFORMED; }

const char* cls_name(int cls) { return cls == HX_DEFORMED ? "deformed" : "cartesian"; }
const char* pde_name(bool visc) { return visc ? "navier_stokes" : "euler"; }
const char* pde_name(const hx_pde* pde)
{
  switch (pde_kind(pde)) {
    case HX_PDE_NAVIER_STOKES: return "navier_stokes";
    case HX_PDE_ADVECTION: return "advection";
    case HX_PDE_SMOOTH_AV: return "smooth_av";
    case HX_PDE_FIX_ADMIS: return "fix_therm_admis";
    default: return "euler";
  }
}

hipEvent_t pool_event(hx_ctx* ctx)
{
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events on the context stream while timing is enabled.
struct Timer
{
  hx_ctx* ctx;
  hipEvent_t a = nullptr, b = nullptr;
  std::string category;
  long long units;
  Timer(hx_ctx* c, std::string cat, long long u) : ctx{c}, category{std::move(cat)}, units{u}
  {
    if (!ctx->timing || ctx->events_off) return;
    a = pool_event(ctx);
    b = pool_event(ctx);
    if (a && b && hipEventRecord(a, ctx->stream) != hipSuccess) a = nullptr;
  }
  ~Timer()
  {
    if (!ctx->timing || ctx->events_off || !a || !b) return;
    if (hipEventRecord(b, ctx->stream) == hipSuccess) ctx->pending.push_back({a, b, category, units});
  }
};

void resolve_timing(hx_ctx* ctx)
{
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      auto& t = ctx->totals[p.category];
      t.ms += ms;
      ++t.launches;
      t.units += p.units;
    }
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
}

__global__ void k_fill(double* p, double v) { p[threadIdx.x] = v; }

// one step's time step on the device, min(Max_dt, max_time_step) * chebyshev_step (Solver.cpp:843-848), with
// Max_dt from *mdt (or 1 with local time stepping), and the status it advances (Solver.cpp:872-876)
__global__ void k_step_dt(const double* mdt, int unit, double max_time_step, double cheby, Step_dev* s, int count)
{
  const double m = unit ? 1. : *mdt;
  double dt = (max_time_step < m ? max_time_step : m) * cheby;
  // the product is rounded before it is accumulated, as the host's flow_time += dt does (no fused multiply-add)
  asm volatile("" : "+v"(dt));
  s->dt = dt;
  s->time_step = dt;
  s->flow_time += dt;
  if (count) ++s->iteration;
}

__global__ void k_set_status(Step_dev* s, double time_step, double flow_time, long long iteration)
{
  s->dt = time_step;
  s->time_step = time_step;
  s->flow_time = flow_time;
  s->iteration = iteration;
}

// batch of cases: per-case time step min(Max_dt, max_time_step) * chebyshev_step (Solver.cpp:843-848), and the time
// step scale of every point set to its case's time step
__global__ __launch_bounds__(256) void k_case_dt(const unsigned long long* __restrict__ case_min, double* __restrict__ case_dt,
                                                 int n_cases, double max_time_step, double cheby)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_cases) return;
  const double mdt = __longlong_as_double((long long)case_min[i]);
  case_dt[i] = (max_time_step < mdt ? max_time_step : mdt) * cheby;
}

// per-case minimum of the time step scale field (which Max_dt in local mode filled with each point's 1/scale) over
// each case's contiguous range of elements, `split` workgroups per case, one atomic minimum per workgroup
__global__ __launch_bounds__(256) void k_case_min_tss(const double* __restrict__ tss, const int* __restrict__ case_start,
                                                      int nq, int split, unsigned long long* __restrict__ case_min)
{
  __shared__ double w[4];
  const int k = blockIdx.x / split, sl = blockIdx.x % split;
  const long long lo = (long long)case_start[k] * nq, hi = (long long)case_start[k + 1] * nq;
  const long long per = (hi - lo + split - 1) / split;
  const long long a = lo + sl * per, b = a + per < hi ? a + per : hi;
  double m = DBL_MAX;
  for (long long i = a + threadIdx.x; i < b; i += 256) m = tss[i] < m ? tss[i] : m;  // NaN never wins
  for (int off = 32; off > 0; off /= 2) {
    const double x = __shfl_xor(m, off, 64);
    m = x < m ? x : m;
  }
  if ((threadIdx.x & 63) == 0) w[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 1; j < 4; ++j) m = w[j] < m ? w[j] : m;
    if (m < DBL_MAX) atomicMin(case_min + k, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ __launch_bounds__(256) void k_case_tss(double* __restrict__ tss, const int* __restrict__ case_id,
                                                  const double* __restrict__ case_dt, int nq, size_t n)
{
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
    tss[i] = case_dt[case_id[i / nq]];
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, unsigned long long v, int n)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// minimum of v[0..n) by atomic minimum into *out (prefilled with DBL_MAX), one atomic per workgroup
__global__ __launch_bounds__(256) void k_min_reduce(const double* __restrict__ v, long long n, double* out)
{
  __shared__ double w[4];
  double m = DBL_MAX;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) m = v[i] < m ? v[i] : m;
  for (int off = 32; off > 0; off /= 2) {
    const double x = __shfl_xor(m, off, 64);
    m = x < m ? x : m;
  }
  if ((threadIdx.x & 63) == 0) w[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) m = w[k] < m ? w[k] : m;
    if (m < DBL_MAX) atomicMin(reinterpret_cast<unsigned long long*>(out), (unsigned long long)__double_as_longlong(m));
  }
}

// plain streaming copy: each workgroup moves chunks of 4 x 256 16-byte items, every lane with its four loads in flight
// before its four stores
constexpr int copy_unroll = 4;
__global__ __launch_bounds__(256) void k_stream_copy(const double2* __restrict__ src, double2* __restrict__ dst, size_t n)
{
  const size_t step = size_t(gridDim.x) * 256 * copy_unroll;
  for (size_t base = size_t(blockIdx.x) * 256 * copy_unroll + threadIdx.x; base < n; base += step) {
    double2 v[copy_unroll];
#pragma unroll
    for (int k = 0; k < copy_unroll; ++k) {
      const size_t i = base + size_t(k) * 256;
      if (i < n) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < copy_unroll; ++k) {
      const size_t i = base + size_t(k) * 256;
      if (i < n) dst[i] = v[k];
    }
  }
}

}  // namespace

namespace {

// vertex sharing by minimum (Solver.cpp:40-54 with Vertex::vector_min): each element copy pushes into its shared
// vertex with a compare-and-swap minimum on the double, then every copy fetches the shared value
__device__ inline void atomic_min_double(double* addr, double v)
{
  unsigned long long* a = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long old = *a;
  while (v < __longlong_as_double(old)) {
    const unsigned long long prev = atomicCAS(a, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

__global__ __launch_bounds__(256) void k_vertex_push_min(const int* vid, const double* vtss, long long n, double* shared)
{
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid < n) atomic_min_double(shared + vid[gid], vtss[gid]);
}

// vertex sharing by maximum (Vertex::vector_max): std::max(shared, v) keeps the shared value unless it is below v
__device__ inline void atomic_max_double(double* addr, double v)
{
  unsigned long long* a = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long old = *a;
  while (__longlong_as_double(old) < v) {
    const unsigned long long prev = atomicCAS(a, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

__global__ __launch_bounds__(256) void k_vertex_push_max(const int* vid, const double* vals, long long n, double* shared)
{
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid < n) atomic_max_double(shared + vid[gid], vals[gid]);
}

__global__ __launch_bounds__(256) void k_vertex_fetch(const int* vid, double* vtss, long long n, const double* shared)
{
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid < n) vtss[gid] = shared[vid[gid]];
}

__global__ __launch_bounds__(256) void k_fill_n(double* p, double v, long long n)
{
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid < n) p[gid] = v;
}

// dst[i] = src[i * stride], i < n
__global__ __launch_bounds__(256) void k_gather_stride(const double* src, int stride, long long n, double* dst)
{
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid < n) dst[gid] = src[gid * stride];
}

// device copies of a call's inputs, freed when the call returns
struct Scratch
{
  std::vector<void*> ptrs;
  ~Scratch()
  {
    // kernels queued before an early error return may still read these buffers
    if (!ptrs.empty()) (void)hipDeviceSynchronize();
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t put(T** dst, const T* src, size_t count)
  {
    *dst = nullptr;
    if (!count) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    *dst = static_cast<T*>(p);
    return src ? hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice) : hipSuccess;
  }
};

// hx_local; project_av_rs > 0 (Advection only): its update also forms the Legendre projection of degree
// project_av_rs - 1 into forcing[0] (the projection pass of update_art_visc_smoothness)
hx_status local_impl(hx_mesh* mesh, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual, int use_filter,
                     int project_av_rs)
{
  if (!live(mesh) || !valid_cls(cls)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  const int kind = pde_kind(pde);
  // HEXED_ASSERTs of the Local constructor (Spatial.hpp:322-323)
  if ((visc || kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS) && i_stage)
    return fail(mesh->ctx, HX_ERR_INVALID, "two-stage stabilization is not applicable to diffusion equations");
  if (i_stage && compute_residual)
    return fail(mesh->ctx, HX_ERR_INVALID, "residual calculation is a single-stage operation");
  const double coef = i_stage ? dt * hx_basis_step_ratio(mesh->dm.rs) : dt;
  if (is_aux(kind) && use_filter)
    return fail(mesh->ctx, HX_ERR_UNSUPPORTED, "the artificial-viscosity and admissibility PDEs are built without the modal filter");
  if ((kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS) && compute_residual)
    return fail(mesh->ctx, HX_ERR_UNSUPPORTED, "the diffusion-only PDEs are built without compute_residual");
  Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/local/" + pde_name(pde), mesh->dm.n_elem[cls]);
  if (kind == HX_PDE_ADVECTION)
    HX_TRY(mesh->ctx, mesh->ks->adv_local(l, cls, coef, i_stage, compute_residual, pde->advect_length, project_av_rs), "local");
  else if (kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS)
    HX_TRY(mesh->ctx, mesh->ks->diff_local(l, cls, kind, coef, pde->diff_time), "local");
  else HX_TRY(mesh->ctx, mesh->ks->local(l, cls, visc, coef, i_stage, compute_residual, use_filter), "local");
  return HX_OK;
}

// compute_advection (kernels_convective.cpp:19, the COMPUTE_CONVECTION sequence); project_av_rs > 0: as local_impl
hx_status compute_advection_impl(hx_mesh* mesh, const hx_kernel_options* o, double advect_length, int project_av_rs)
{
  hx_pde p{};
  p.kind = HX_PDE_ADVECTION;
  p.advect_length = advect_length;
  hx_status s;
  if ((s = hx_neighbor(mesh, HX_CARTESIAN, &p, o->i_stage))) return s;
  if ((s = hx_neighbor(mesh, HX_DEFORMED, &p, o->i_stage))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 0))) return s;
  for (int c = 0; c < 2; ++c)
    if ((s = local_impl(mesh, c, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter, project_av_rs))) return s;
  return hx_prolong_refined(mesh, &p, 0, 0);
}

}  // namespace

extern "C" {

int hx_kernels_built(int n_dim, int row_size, int pde_kind)
{
  if (pde_kind != HX_PDE_EULER && pde_kind != HX_PDE_NAVIER_STOKES && !is_aux(pde_kind)) return 0;
  return kernel_set(n_dim, row_size) ? 1 : 0;
}

hx_status hx_ctx_create(int device, hx_ctx** out)
{
  if (!out) return HX_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return HX_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return HX_ERR_INVALID;
  auto* ctx = new hx_ctx;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&ctx->scalar, 8 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&ctx->host_scalar, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    if (ctx->scalar) (void)hipFree(ctx->scalar);
    delete ctx;
    return HX_ERR_HIP;
  }
  // the scratch array of hx_update_no_alloc, with this device current; a failure is reported by that entry point
  ctx->scratch = scratch_on(device, &ctx->scratch_err);
  *out = ctx;
  return HX_OK;
}

hx_status hx_ctx_destroy(hx_ctx* ctx)
{
  if (!ctx) return HX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  // meshes still alive are detached: their device memory is freed now, their handles stay valid for
  // hx_mesh_destroy, and every other call on them fails with HX_ERR_INVALID
  for (hx_group* g : ctx->groups) release_group(g);
  ctx->groups.clear();
  for (hx_mesh* mesh : ctx->meshes) {
    release_mesh_memory(mesh);
    mesh->ctx = nullptr;
  }
  ctx->meshes.clear();
  for (void* p : ctx->ops)
    if (p) (void)hipFree(p);
  for (void* p : ctx->eops)
    if (p) (void)hipFree(p);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipFree(ctx->scalar);
  (void)hipHostFree(ctx->host_scalar);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HX_OK;
}

const char* hx_last_error(const hx_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

hx_status hx_ctx_synchronize(hx_ctx* ctx)
{
  if (!ctx) return HX_ERR_INVALID;
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return HX_OK;
}

void* hx_ctx_stream(hx_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

hx_status hx_stream_copy(hx_ctx* ctx, size_t bytes, int reps, double* gbs)
{
  if (!ctx || !gbs || reps < 1 || bytes < 16) return HX_ERR_INVALID;
  const size_t n = bytes / 16;
  void* a = nullptr;
  void* b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hx_status st = HX_OK;
  auto cleanup = [&]() {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  if ((st = hip_check(ctx, hipMalloc(&a, n * 16), "hipMalloc(copy)")) || (st = hip_check(ctx, hipMalloc(&b, n * 16), "hipMalloc(copy)")) ||
      (st = hip_check(ctx, hipMemsetAsync(a, 0, n * 16, ctx->stream), "memset(copy)")) ||
      (st = hip_check(ctx, hipEventCreate(&e0), "event")) || (st = hip_check(ctx, hipEventCreate(&e1), "event"))) {
    cleanup();
    return st;
  }
  // the best of several grid sizes: 8, 32 and 128 workgroups per compute unit, and one workgroup per chunk
  const size_t chunks = (n + 256 * copy_unroll - 1) / (256 * copy_unroll);
  double best = 0;
  for (size_t cap : {size_t(256) * 8, size_t(256) * 32, size_t(256) * 128, chunks}) {
    const unsigned grid = unsigned(std::max<size_t>(1, std::min(chunks, cap)));
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, (const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, (const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e1, ctx->stream);
    if ((st = hip_check(ctx, hipEventSynchronize(e1), "copy sync"))) {
      cleanup();
      return st;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    best = std::max(best, 2.0 * double(n * 16) * reps / (ms * 1e-3) / 1e9);
  }
  *gbs = best;
  cleanup();
  return HX_OK;
}

hx_status hx_ctx_set_timing(hx_ctx* ctx, int enable)
{
  if (!ctx) return HX_ERR_INVALID;
  ctx->timing = enable != 0;
  return HX_OK;
}

hx_status hx_ctx_timing_reset(hx_ctx* ctx)
{
  if (!ctx) return HX_ERR_INVALID;
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "timing sync");
  resolve_timing(ctx);
  ctx->totals.clear();
  return HX_OK;
}

hx_status hx_ctx_timing(hx_ctx* ctx, const char* category, double* total_ms, long long* launches, long long* units)
{
  if (!ctx || !category) return HX_ERR_INVALID;
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "timing sync");
  resolve_timing(ctx);
  auto it = ctx->totals.find(category);
  const hx_ctx::Total t = it == ctx->totals.end() ? hx_ctx::Total{} : it->second;
  if (total_ms) *total_ms = t.ms;
  if (launches) *launches = t.launches;
  if (units) *units = t.units;
  return HX_OK;
}

hx_status hx_ctx_timing_list(hx_ctx* ctx, char* out, size_t capacity)
{
  if (!ctx || !out || capacity == 0) return HX_ERR_INVALID;
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "timing sync");
  resolve_timing(ctx);
  std::string all;
  for (auto& kv : ctx->totals) all += kv.first + "\n";
  std::snprintf(out, capacity, "%s", all.c_str());
  return all.size() < capacity ? HX_OK : HX_ERR_INVALID;
}

int hx_basis_table(int row_size, const char* name, double* out, int n)
{
  using namespace hx_basis_tables;
  if (row_size < HX_MIN_ROW_SIZE || row_size > HX_MAX_ROW_SIZE || !name) return -1;
  const int k = row_size - 2, rs = row_size;
  const double* src = nullptr;
  int size = 0;
  std::string s(name);
  if (s == "node") { src = nodes[k]; size = rs; }
  else if (s == "weight") { src = weights[k]; size = rs; }
  else if (s == "diff_mat") { src = diff_mats[k]; size = rs * rs; }
  else if (s == "boundary") { src = boundarys[k]; size = 2 * rs; }
  else if (s == "orthogonal") { src = orthogonals[k]; size = rs * rs; }
  else if (s == "filter") { src = filters[k]; size = rs * rs; }
  else if (s == "prolong") { src = prolongs[k]; size = 2 * rs * rs; }
  else if (s == "restrict") { src = restricts[k]; size = 2 * rs * rs; }
  else if (s == "eigenvals") { src = eigenvalss[k]; size = 2; }
  else return -1;
  if (out) std::memcpy(out, src, sizeof(double) * std::min(std::max(n, 0), size));
  return size;
}

double hx_basis_max_cfl(int row_size)
{
  if (row_size < HX_MIN_ROW_SIZE || row_size > HX_MAX_ROW_SIZE) return NAN;
  return -2 * 0.9 / hx_basis_tables::eigenvalss[row_size - 2][0];
}

double hx_basis_step_ratio(int) { return .5 / 0.9; }

hx_status hx_mesh_create(hx_ctx* ctx, const hx_host_mesh* d, hx_mesh** out)
{
  if (!ctx || !d || !out) return HX_ERR_INVALID;
  *out = nullptr;
  if (d->n_dim < 1 || d->n_dim > HX_MAX_N_DIM || d->row_size < HX_MIN_ROW_SIZE || d->row_size > HX_MAX_ROW_SIZE)
    return fail(ctx, HX_ERR_INVALID, "demand for invalid kernel");
  const Kernel_set* ks = kernel_set(d->n_dim, d->row_size);
  if (!ks) return fail(ctx, HX_ERR_UNSUPPORTED, "kernels for this n_dim/row_size are not built");
  Sizes z(d->n_dim, d->row_size);
  std::string bad = validate(d, z);
  if (!bad.empty()) return fail(ctx, HX_ERR_INVALID, "invalid mesh: " + bad);
  HX_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  hx_status s = ensure_ops(ctx, d->row_size);
  if (s != HX_OK) return s;
  auto* mesh = new hx_mesh;
  mesh->ctx = ctx;
  mesh->ks = ks;
  Dev_mesh& m = mesh->dm;
  m.nd = z.nd; m.rs = z.rs; m.nq = z.nq; m.nfq = z.nfq; m.nv = z.nv; m.face_sz = z.face_sz;
  auto bail = [&](hx_status st) {
    hx_mesh_destroy(mesh);
    return st;
  };
  for (int c = 0; c < 2; ++c) {
    const long long n = d->n_elem[c];
    m.n_elem[c] = int(n);
    if ((s = device_copy(mesh, &m.state[c], d->state[c], n * z.nv * z.nq, true))) return bail(s);
    if ((s = device_copy(mesh, &m.rcache[c], d->rcache[c], n * std::max(z.nv, z.rs) * z.nq, true))) return bail(s);
    if ((s = device_copy(mesh, &m.tss[c], d->tss[c], n * z.nq, true))) return bail(s);
    if ((s = device_copy(mesh, &m.av[c], d->av[c], n * 2 * z.nq, false))) return bail(s);
    if ((s = device_copy(mesh, &m.forcing[c], d->forcing[c], n * 4 * z.nq, false))) return bail(s);
    if ((s = device_copy(mesh, &m.advection[c], d->advection[c], n * z.rs * z.nq, false))) return bail(s);
    if ((s = device_copy(mesh, &m.nom_size[c], d->nom_size[c], n, true))) return bail(s);
    if ((s = device_copy(mesh, &m.vertex_tss[c], d->vertex_tss[c], n * z.nvert, true))) return bail(s);
    if ((s = device_copy(mesh, &m.face_slot[c], d->face_slot[c], n * 2 * z.nd, true))) return bail(s);
    m.n_con[c] = d->n_con[c];
    if ((s = device_copy(mesh, &m.con_slot[c], d->con_slot[c], 2LL * d->n_con[c], true))) return bail(s);
    if ((s = device_copy(mesh, &m.con_dir[c], d->con_dir[c], (long long)d->n_con[c], true))) return bail(s);
  }
  const long long n_def = d->n_elem[HX_DEFORMED];
  if ((s = device_copy(mesh, &m.metrics, d->metrics, n_def * (z.nd * z.nd + 1) * z.nq, true))) return bail(s);
  if ((s = device_copy(mesh, &m.face_nslot, d->face_nslot, n_def * 2 * z.nd, true))) return bail(s);
  m.n_slot = d->n_slot;
  if ((s = device_copy(mesh, &m.faces, d->faces, (long long)d->n_slot * z.face_sz, true))) return bail(s);
  if ((s = device_copy(mesh, &m.faces_alt, (const double*)nullptr, (long long)d->n_slot * z.nv * z.nfq, true))) return bail(s);
  m.n_nslot = d->n_nslot;
  if ((s = device_copy(mesh, &m.normals, d->normals, (long long)d->n_nslot * z.nd * z.nfq, true))) return bail(s);
  if ((s = device_copy(mesh, &m.con_nslot, d->con_nslot, 2LL * d->n_con[HX_DEFORMED], true))) return bail(s);
  m.n_ref = d->n_ref;
  if ((s = device_copy(mesh, &m.ref_slot, d->ref_slot, 5LL * d->n_ref, true))) return bail(s);
  if ((s = device_copy(mesh, &m.ref_stretch, d->ref_stretch, 2LL * d->n_ref, true))) return bail(s);
  m.n_bc_con = d->n_bc_con;
  if ((s = device_copy(mesh, &m.bc_con, d->bc_con, (long long)d->n_bc_con, true))) return bail(s);
  if ((s = device_copy(mesh, &m.bc_serial, d->bc_serial, (long long)d->n_bc_con, true))) return bail(s);
  m.n_bc = d->n_bc;
  if ((s = device_copy(mesh, &m.bcs, d->bcs, (long long)d->n_bc, true))) return bail(s);
  if ((s = device_copy(mesh, &m.bc_cache, d->bc_cache, (long long)d->n_bc_con * z.nv * z.nfq, true))) return bail(s);
  {
    std::vector<int> partner[2], mortar[2], ref_con;
    mesh->mortar_fusable = build_partners(d, z, partner, mortar, ref_con);
    if ((s = device_copy(mesh, &m.ref_con, ref_con.data(), (long long)ref_con.size(), true))) return bail(s);
    for (int c = 0; c < 2; ++c) {
      if ((s = device_copy(mesh, &m.partner[c], partner[c].data(), (long long)partner[c].size(), true))) return bail(s);
      m.n_mortar_con[c] = int(mortar[c].size());
      if ((s = device_copy(mesh, &m.mortar_con[c], mortar[c].data(), (long long)mortar[c].size(), true))) return bail(s);
    }
  }
  ctx->meshes.push_back(mesh);
  *out = mesh;
  return HX_OK;
}

hx_status hx_mesh_validate(const hx_host_mesh* d, char* msg, size_t capacity)
{
  std::string bad;
  if (!d) bad = "no mesh";
  else if (d->n_dim < 1 || d->n_dim > HX_MAX_N_DIM || d->row_size < HX_MIN_ROW_SIZE || d->row_size > HX_MAX_ROW_SIZE)
    bad = "demand for invalid kernel";
  else {
    std::string why = validate(d, Sizes(d->n_dim, d->row_size));
    if (!why.empty()) bad = "invalid mesh: " + why;
  }
  if (msg && capacity) std::snprintf(msg, capacity, "%s", bad.c_str());
  return bad.empty() ? HX_OK : HX_ERR_INVALID;
}

hx_status hx_mesh_destroy(hx_mesh* mesh)
{
  if (!mesh) return HX_OK;
  if (mesh->n_groups) return fail(mesh->ctx, HX_ERR_INVALID, "the mesh belongs to a group: destroy the group first");
  if (hx_ctx* ctx = mesh->ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_mesh_memory(mesh);
    ctx->meshes.erase(std::remove(ctx->meshes.begin(), ctx->meshes.end(), mesh), ctx->meshes.end());
  }
  // a mesh detached by hx_ctx_destroy holds no device memory any more
  delete mesh;
  return HX_OK;
}

hx_status hx_mesh_upload(hx_mesh* mesh, int field, int cls, const double* host, size_t n)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  ++mesh->gen;
  double** p = field_ptr(mesh->dm, field, cls);
  long long sz = field_size(mesh->dm, field, cls);
  if (!p || sz < 0) return fail(mesh->ctx, HX_ERR_INVALID, "unknown field or class");
  if ((long long)n != sz) return fail(mesh->ctx, HX_ERR_INVALID, "field size mismatch");
  if (sz == 0) return HX_OK;
  if (!*p) {
    // optional fields (artificial viscosity, forcing, advection) are allocated on first upload
    void* q = nullptr;
    HX_TRY(mesh->ctx, hipMalloc(&q, n * sizeof(double)), "hipMalloc(field)");
    mesh->allocs.push_back(q);
    *p = static_cast<double*>(q);
  }
  HX_TRY(mesh->ctx, hipMemcpyAsync(*p, host, n * sizeof(double), hipMemcpyHostToDevice, mesh->ctx->stream), "upload");
  HX_TRY(mesh->ctx, hipStreamSynchronize(mesh->ctx->stream), "upload sync");
  return HX_OK;
}

hx_status hx_mesh_download(hx_mesh* mesh, int field, int cls, double* host, size_t n)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  double** p = field_ptr(mesh->dm, field, cls);
  long long sz = field_size(mesh->dm, field, cls);
  if (!p || sz < 0) return fail(mesh->ctx, HX_ERR_INVALID, "unknown field or class");
  if ((long long)n != sz) return fail(mesh->ctx, HX_ERR_INVALID, "field size mismatch");
  if (sz == 0) return HX_OK;
  if (!*p) {
    std::memset(host, 0, n * sizeof(double));
    return HX_OK;
  }
  HX_TRY(mesh->ctx, hipMemcpyAsync(host, *p, n * sizeof(double), hipMemcpyDeviceToHost, mesh->ctx->stream), "download");
  HX_TRY(mesh->ctx, hipStreamSynchronize(mesh->ctx->stream), "download sync");
  return HX_OK;
}

hx_status hx_mesh_set_precision(hx_mesh* mesh, int precision)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  if (precision != HX_FP64 && precision != HX_FP32_MFMA) return fail(mesh->ctx, HX_ERR_INVALID, "unknown precision");
  ++mesh->gen;
  mesh->precision = precision;
  return HX_OK;
}

hx_status hx_mesh_device_ptr(hx_mesh* mesh, int field, int cls, void** out, size_t* n)
{
  if (!live(mesh) || !out || !n) return HX_ERR_INVALID;
  ++mesh->gen;  // the caller may write through the pointer
  double** p = field_ptr(mesh->dm, field, cls);
  long long sz = field_size(mesh->dm, field, cls);
  if (!p || sz < 0) return fail(mesh->ctx, HX_ERR_INVALID, "unknown field or class");
  *out = *p;
  *n = size_t(sz);
  return HX_OK;
}

hx_status hx_mesh_set_face_fold(hx_mesh* mesh, int mode)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  if (mode < -1 || mode > 1) return fail(mesh->ctx, HX_ERR_INVALID, "face fold mode: -1, 0 or 1");
  // a graph captured by hx_update_async holds the launches of the previous mode
  if (mode != mesh->face_fold_mode) release_graph(mesh);
  mesh->face_fold_mode = mode;
  return HX_OK;
}

hx_status hx_mesh_mark_changed(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  ++mesh->gen;
  return HX_OK;
}

hx_status hx_write_face(hx_mesh* mesh, int cls, const hx_pde* pde)
{
  if (!live(mesh) || !valid_cls(cls)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/write face/" + pde_name(pde), mesh->dm.n_elem[cls]);
  if (is_aux(pde_kind(pde))) HX_TRY(mesh->ctx, mesh->ks->av_write_face(l, cls, pde->kind), "write_face");
  else HX_TRY(mesh->ctx, mesh->ks->write_face(l, cls, 0), "write_face");
  return HX_OK;
}

hx_status hx_neighbor(hx_mesh* mesh, int cls, const hx_pde* pde, int)
{
  if (!live(mesh) || !valid_cls(cls)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/neighbor/" + pde_name(pde), mesh->dm.n_con[cls]);
  if (is_aux(pde_kind(pde))) HX_TRY(mesh->ctx, mesh->ks->av_neighbor(l, cls, pde->kind), "neighbor");
  else HX_TRY(mesh->ctx, mesh->ks->neighbor(l, cls, visc), "neighbor");
  return HX_OK;
}

hx_status hx_local(hx_mesh* mesh, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual, int use_filter)
{
  return local_impl(mesh, cls, pde, dt, i_stage, compute_residual, use_filter, 0);
}

hx_status hx_neighbor_reconcile(hx_mesh* mesh, int cls, const hx_pde* pde)
{
  if (!live(mesh) || !valid_cls(cls)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  const int kind = pde_kind(pde);
  const bool diff_only = kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS;
  Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/neighbor reconcile/" + (diff_only ? pde_name(pde) : "navier_stokes"),
           mesh->dm.n_con[cls]);
  if (diff_only) HX_TRY(mesh->ctx, mesh->ks->diff_neighbor_reconcile(l, cls, kind), "neighbor_reconcile");
  else HX_TRY(mesh->ctx, mesh->ks->neighbor_reconcile(l, cls), "neighbor_reconcile");
  return HX_OK;
}

hx_status hx_reconcile_ldg_flux(hx_mesh* mesh, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual,
                                int use_filter)
{
  if (!live(mesh) || !valid_cls(cls)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  const int kind = pde_kind(pde);
  if ((visc || kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS) && i_stage)
    return fail(mesh->ctx, HX_ERR_INVALID, "two-stage stabilization is not applicable to diffusion equations");
  if (kind == HX_PDE_ADVECTION)
    return fail(mesh->ctx, HX_ERR_INVALID, "Reconcile_ldg_flux needs a PDE with diffusion");
  if (kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS) {
    if (compute_residual || use_filter)
      return fail(mesh->ctx, HX_ERR_UNSUPPORTED, "the diffusion-only PDEs are built without compute_residual and the filter");
    Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/reconcile LDG flux/" + pde_name(pde), mesh->dm.n_elem[cls]);
    HX_TRY(mesh->ctx, mesh->ks->diff_reconcile(l, cls, kind, dt, pde->diff_time, pde->cheby_step), "reconcile_ldg_flux");
    return HX_OK;
  }
  Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/reconcile LDG flux/navier_stokes", mesh->dm.n_elem[cls]);
  HX_TRY(mesh->ctx, mesh->ks->reconcile_ldg_flux(l, cls, dt, compute_residual, use_filter), "reconcile_ldg_flux");
  return HX_OK;
}

hx_status hx_max_dt(hx_mesh* mesh, int cls, const hx_pde* pde, int is_local, int, double safety_conv, double safety_diff,
                    double* dt_out)
{
  if (!live(mesh) || !valid_cls(cls) || !dt_out) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  const double max_cfl_c = hx_basis_max_cfl(mesh->dm.rs) * safety_conv;
  const double max_cfl_d = -2 / hx_basis_tables::eigenvalss[mesh->dm.rs - 2][1] * safety_diff;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, l.stream, l.scalar, DBL_MAX);
  HX_TRY(mesh->ctx, hipGetLastError(), "fill");
  {
    Timer tm(mesh->ctx, std::string(cls_name(cls)) + "/compute time step/" + pde_name(pde), mesh->dm.n_elem[cls]);
    // Smooth_art_visc and Fix_therm_admis share Max_dt's diffusivity-1 form (pde.hpp:412-416, 486-490)
    if (is_aux(pde_kind(pde)))
      HX_TRY(mesh->ctx, mesh->ks->av_max_dt(l, cls, pde->kind != HX_PDE_ADVECTION, is_local, max_cfl_c, max_cfl_d),
             "max_dt");
    else HX_TRY(mesh->ctx, mesh->ks->max_dt(l, cls, visc, is_local, max_cfl_c, max_cfl_d), "max_dt");
  }
  HX_TRY(mesh->ctx, hipMemcpyAsync(mesh->ctx->host_scalar, l.scalar, sizeof(double), hipMemcpyDeviceToHost, l.stream),
         "max_dt readback");
  HX_TRY(mesh->ctx, hipStreamSynchronize(l.stream), "max_dt sync");
  const double dt = *mesh->ctx->host_scalar;
  *dt_out = is_local ? 1. : dt;
  return HX_OK;
}

hx_status hx_prolong_refined(hx_mesh* mesh, const hx_pde* pde, int scale, int offset)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  Timer tm(mesh->ctx, "prolong/restrict/prolong", mesh->dm.n_ref);
  HX_TRY(mesh->ctx, mesh->ks->prolong(l, n_extrap(mesh, pde), scale, offset), "prolong");
  return HX_OK;
}

hx_status hx_restrict_refined(hx_mesh* mesh, const hx_pde* pde, int scale, int offset)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Launch l = make_launch(mesh);
  bool visc;
  if (hx_status s = pde_check(mesh, pde, &visc, &l)) return s;
  Timer tm(mesh->ctx, "prolong/restrict/restrict", mesh->dm.n_ref);
  HX_TRY(mesh->ctx, mesh->ks->restrict_(l, n_extrap(mesh, pde), scale, offset), "restrict");
  return HX_OK;
}

hx_status hx_apply_state_bcs(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Timer tm(mesh->ctx, "boundary conditions/state", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->state_bcs(make_launch(mesh)), "state_bcs");
  return HX_OK;
}

hx_status hx_apply_flux_bcs(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Timer tm(mesh->ctx, "boundary conditions/flux", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->flux_bcs(make_launch(mesh)), "flux_bcs");
  return HX_OK;
}

hx_status hx_compute_euler(hx_mesh* mesh, const hx_kernel_options* o)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_EULER;
  hx_status s;
  if (mesh->fold && mesh->dm.nd == 3 && mesh->dm.n_ref > 0 && mesh->mortar_fusable) {
    // the folded stage's Neighbor covers only the mortar connections: one launch with the restriction
    Launch l = make_launch(mesh);
    Timer tm(mesh->ctx, "prolong/restrict/mortar flux and restrict", mesh->dm.n_ref);
    HX_TRY(mesh->ctx, mesh->ks->mortar_restrict(l), "mortar flux and restrict");
  } else {
    if ((s = hx_neighbor(mesh, HX_CARTESIAN, &p, o->i_stage))) return s;
    if ((s = hx_neighbor(mesh, HX_DEFORMED, &p, o->i_stage))) return s;
    if ((s = hx_restrict_refined(mesh, &p, 1, 0))) return s;
  }
  if ((s = hx_local(mesh, HX_CARTESIAN, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_local(mesh, HX_DEFORMED, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
  return hx_prolong_refined(mesh, &p, 0, 0);
}

hx_status hx_compute_navier_stokes(hx_mesh* mesh, const hx_kernel_options* o, hx_transport visc, hx_transport cond)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_NAVIER_STOKES;
  p.visc = visc;
  p.cond = cond;
  if (o->i_stage)
    return fail(mesh->ctx, HX_ERR_INVALID, "two-stage stabilization is not applicable to diffusion equations");
  hx_status s;
  if ((s = hx_neighbor(mesh, HX_CARTESIAN, &p, o->i_stage))) return s;
  if ((s = hx_neighbor(mesh, HX_DEFORMED, &p, o->i_stage))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 0))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 0, 1))) return s;
  if ((s = hx_local(mesh, HX_CARTESIAN, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_local(mesh, HX_DEFORMED, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
  if (!o->i_stage) {
    if ((s = hx_prolong_refined(mesh, &p, 1, 1))) return s;
    if ((s = hx_apply_flux_bcs(mesh))) return s;
    if ((s = hx_neighbor_reconcile(mesh, HX_CARTESIAN, &p))) return s;
    if ((s = hx_neighbor_reconcile(mesh, HX_DEFORMED, &p))) return s;
    if ((s = hx_restrict_refined(mesh, &p, 1, 1))) return s;
    if ((s = hx_reconcile_ldg_flux(mesh, HX_CARTESIAN, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
    if ((s = hx_reconcile_ldg_flux(mesh, HX_DEFORMED, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter))) return s;
  }
  return hx_prolong_refined(mesh, &p, 0, 0);
}

hx_status hx_max_dt_euler(hx_mesh* mesh, const hx_kernel_options* o, double msc, double msd, int local_time, double* dt_out)
{
  if (!live(mesh) || !dt_out) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_EULER;
  double a, b;
  hx_status s;
  const int uf = o ? o->use_filter : 0;
  if ((s = hx_max_dt(mesh, HX_CARTESIAN, &p, local_time, uf, msc, msd, &a))) return s;
  if ((s = hx_max_dt(mesh, HX_DEFORMED, &p, local_time, uf, msc, msd, &b))) return s;
  *dt_out = std::min(a, b);
  return HX_OK;
}

hx_status hx_max_dt_navier_stokes(hx_mesh* mesh, const hx_kernel_options* o, double msc, double msd, int local_time,
                                  hx_transport visc, hx_transport cond, double* dt_out)
{
  if (!live(mesh) || !dt_out) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_NAVIER_STOKES;
  p.visc = visc;
  p.cond = cond;
  double a, b;
  hx_status s;
  const int uf = o ? o->use_filter : 0;
  if ((s = hx_max_dt(mesh, HX_CARTESIAN, &p, local_time, uf, msc, msd, &a))) return s;
  if ((s = hx_max_dt(mesh, HX_DEFORMED, &p, local_time, uf, msc, msd, &b))) return s;
  *dt_out = std::min(a, b);
  return HX_OK;
}

hx_status hx_compute_write_face(hx_mesh* mesh)
{
  hx_status s;
  if ((s = hx_write_face(mesh, HX_CARTESIAN, nullptr))) return s;
  return hx_write_face(mesh, HX_DEFORMED, nullptr);
}

hx_status hx_compute_prolong(hx_mesh* mesh, int scale, int offset) { return hx_prolong_refined(mesh, nullptr, scale, offset); }
hx_status hx_compute_restrict(hx_mesh* mesh, int scale, int offset) { return hx_restrict_refined(mesh, nullptr, scale, offset); }

}  // extern "C"

namespace {

// member-wise equality (structure padding is not compared)
bool same_transport(const hx_transport& a, const hx_transport& b)
{
  return a.const_val == b.const_val && a.ref_val == b.ref_val && a.ref_temp == b.ref_temp &&
         a.temp_offset == b.temp_offset && a.is_viscous == b.is_viscous;
}
bool same_options(const hx_update_options& a, const hx_update_options& b)
{
  return a.max_safety == b.max_safety && a.max_time_step == b.max_time_step && a.n_cheby_flow == b.n_cheby_flow &&
         a.flow_iters == b.flow_iters && a.local_time == b.local_time && a.use_filter == b.use_filter &&
         a.use_ldg == b.use_ldg && same_transport(a.visc, b.visc) && same_transport(a.cond, b.cond) &&
         a.fix_admis == b.fix_admis && a.fix_admis_max_safety == b.fix_admis_max_safety;
}

// math::chebyshev_step (math.cpp:104-107)
double cheby_step(double n, double i) { return 1 / (1 - std::cos((n - i - 0.5) * M_PI / n)); }

// the time-step settings of Solver::update (Solver.cpp:838-849), shared by its host and device forms
struct Update_setup
{
  double msc, msd;   // safety factors of Max_dt (convective over the largest Chebyshev step, diffusive)
  int n_cheby;
  bool fusable;      // with a global time step, stage 1's Local also computes the next step's Max_dt
  Ns_params ns;
  double* block_min; // the fused per-workgroup minima: the mesh's buffer, or the scratch array
  long long n_blocks;
  bool in_scratch;
  double* scalar;    // the reduction scalar
};

// workgroups of the Euler Local over both classes, each storing one fused minimum
long long euler_local_blocks(const hx_mesh* mesh)
{
  const long long epb = std::max(1, 256 / mesh->dm.nq);  // elements per workgroup of the Euler Local
  return (mesh->dm.n_elem[0] + epb - 1) / epb + (mesh->dm.n_elem[1] + epb - 1) / epb;
}

// no_alloc: the minima and the reduction scalar go to the file-scope scratch array, and nothing is allocated
hx_status update_setup(hx_mesh* mesh, const hx_update_options* o, Update_setup* u, bool no_alloc = false)
{
  hx_ctx* ctx = mesh->ctx;
  // the state that stage 1's Local writes is the state that Max_dt would read next; local time stepping writes the
  // time step scale in Max_dt, so it is not fused
  u->fusable = !o->local_time;
  u->n_blocks = euler_local_blocks(mesh);
  u->in_scratch = no_alloc;
  if (no_alloc) {
    if (u->fusable && u->n_blocks > no_alloc_max_blocks)
      return fail(ctx, HX_ERR_INVALID, "mesh too large for hx_update_no_alloc: " + std::to_string(u->n_blocks) +
                                           " Euler Local workgroups, at most " + std::to_string(no_alloc_max_blocks));
    u->block_min = ctx->scratch + scratch_minima_offset;
    u->scalar = ctx->scratch;
  } else {
    if (u->fusable && !mesh->dt_block_min) {
      mesh->n_dt_blocks = u->n_blocks;
      void* p = nullptr;
      HX_TRY(ctx, hipMalloc(&p, std::max<long long>(mesh->n_dt_blocks, 1) * sizeof(double)), "hipMalloc(time step minima)");
      mesh->allocs.push_back(p);
      mesh->dt_block_min = static_cast<double*>(p);
    }
    u->block_min = mesh->dt_block_min;
    u->scalar = ctx->scalar;
  }
  u->n_cheby = o->n_cheby_flow;
  const double max_cheby = cheby_step(u->n_cheby, u->n_cheby - 1);
  u->msc = o->max_safety / max_cheby;
  u->msd = o->max_safety;
  u->ns = Ns_params{};
  if (o->use_ldg) {
    hx_pde p{};
    p.kind = HX_PDE_NAVIER_STOKES;
    p.visc = o->visc;
    p.cond = o->cond;
    Launch lp{};
    bool visc;
    if (hx_status s = pde_check(mesh, &p, &visc, &lp)) return s;
    u->ns = lp.ns;
  }
  return HX_OK;
}

// whether the previous update's fused Max_dt minima are still those of the current state and settings
bool fused_pending(hx_mesh* mesh, const hx_update_options* o, const Update_setup& u)
{
  const auto& c = mesh->dt_cache;
  return u.fusable && c.valid && c.gen == mesh->gen && c.msc == u.msc && c.msd == u.msd && c.use_ldg == o->use_ldg &&
         same_transport(c.visc, o->visc) && same_transport(c.cond, o->cond) && c.in_scratch == u.in_scratch &&
         (!u.in_scratch || mesh->scratch_epoch == scratch_epoch(mesh->ctx->device));
}

void remember_fused(hx_mesh* mesh, const hx_update_options* o, const Update_setup& u)
{
  auto& c = mesh->dt_cache;
  c.valid = true;
  c.gen = mesh->gen;
  c.msc = u.msc;
  c.msd = u.msd;
  c.use_ldg = o->use_ldg;
  c.visc = o->visc;
  c.cond = o->cond;
  c.in_scratch = u.in_scratch;
}

// Max_dt of both classes into the context's reduction scalar, without reading it back (kernels_max_dt.cpp:8-17)
hx_status enqueue_max_dt(hx_mesh* mesh, const hx_update_options* o, const Update_setup& u)
{
  hx_ctx* ctx = mesh->ctx;
  const double max_cfl_c = hx_basis_max_cfl(mesh->dm.rs) * u.msc;
  const double max_cfl_d = -2 / hx_basis_tables::eigenvalss[mesh->dm.rs - 2][1] * u.msd;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, ctx->stream, u.scalar, DBL_MAX);
  HX_TRY(ctx, hipGetLastError(), "fill");
  for (int cls = 0; cls < 2; ++cls) {
    Launch l = make_launch(mesh);
    l.ns = u.ns;
    Timer tm(ctx, std::string(cls_name(cls)) + "/compute time step/" + pde_name(o->use_ldg != 0), mesh->dm.n_elem[cls]);
    HX_TRY(ctx, mesh->ks->max_dt(l, cls, o->use_ldg != 0, o->local_time, max_cfl_c, max_cfl_d), "max_dt");
  }
  return HX_OK;
}

// the fused minima of the previous stage 1 reduced into the context's reduction scalar
hx_status enqueue_fused_reduce(hx_mesh* mesh, const Update_setup& u)
{
  hx_ctx* ctx = mesh->ctx;
  Timer tm(ctx, "fused/compute time step/reduce", u.n_blocks);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, ctx->stream, u.scalar, DBL_MAX);
  const unsigned g = unsigned(std::min<long long>(256, (u.n_blocks + 255) / 256));
  if (g) hipLaunchKernelGGL(k_min_reduce, dim3(g), dim3(256), 0, ctx->stream, u.block_min, u.n_blocks, u.scalar);
  HX_TRY(ctx, hipGetLastError(), "time step reduction");
  return HX_OK;
}

// Whether hx_update folds the face fluxes into the Local: by default where an element has no more face points than
// volume points (2 n_dim rs^(n_dim-1) <= rs^n_dim, i.e. rs >= 2 n_dim), where the neighbour's face states the Local
// then reads cost less than the Neighbor launch saves (measured: 3D row size 6 folded is 20-28% faster per step, row
// size 4 on a tree mesh 8% slower)
bool folds_faces(const hx_mesh* mesh)
{
  return mesh->face_fold_mode == 1 || (mesh->face_fold_mode == -1 && mesh->dm.rs >= 2 * mesh->dm.nd);
}

// the face-flux folding flags of one stage of an update, reset when the stage ends
struct Stage_fold
{
  hx_mesh* mesh;
  Stage_fold(hx_mesh* m, bool on, int stage) : mesh{m}
  {
    mesh->fold = on ? 1 : 0;
    mesh->alt_rd = on ? stage : 0;
    mesh->alt_wr = on ? 1 - stage : 0;
  }
  ~Stage_fold() { mesh->fold = mesh->alt_rd = mesh->alt_wr = 0; }
};

// The kernels of one Solver::update (Solver.cpp:834-886), with fix_admissibility after each stage when enabled
// (Solver.cpp:867). Host form (st not null): each step's Max_dt is read back and the time step passed by value.
// Device form (st null): the time step stays in mesh->step, formed by k_step_dt, and the stage kernels read it; nothing
// here waits for the device. `pending`: the previous stage 1's fused Max_dt minima hold the current state's.
hx_status update_once(hx_mesh* mesh, const hx_update_options* o, const Update_setup& u, hx_update_status* st, bool& pending)
{
  hx_ctx* ctx = mesh->ctx;
  const bool device_dt = st == nullptr;
  for (int i_flow = 0; i_flow < o->flow_iters; ++i_flow) {
    for (int i_cheby = 0; i_cheby < u.n_cheby; ++i_cheby) {
      hx_status s;
      double dt = 1.;
      if (pending) {
        if ((s = enqueue_fused_reduce(mesh, u))) return s;
      } else if (device_dt) {
        if ((s = enqueue_max_dt(mesh, o, u))) return s;
      } else {
        hx_kernel_options ko{0., 0, 0, o->use_filter};
        double mdt;
        s = o->use_ldg ? hx_max_dt_navier_stokes(mesh, &ko, u.msc, u.msd, o->local_time, o->visc, o->cond, &mdt)
                       : hx_max_dt_euler(mesh, &ko, u.msc, u.msd, o->local_time, &mdt);
        if (s) return s;
        dt = std::min(mdt, o->max_time_step) * cheby_step(u.n_cheby, i_cheby);
      }
      if (device_dt) {
        hipLaunchKernelGGL(k_step_dt, dim3(1), dim3(1), 0, ctx->stream, u.scalar, pending ? 0 : o->local_time,
                           o->max_time_step, cheby_step(u.n_cheby, i_cheby), mesh->step, int(i_flow == 0 && i_cheby == 0));
        HX_TRY(ctx, hipGetLastError(), "time step");
      } else if (pending) {
        HX_TRY(ctx, hipMemcpyAsync(ctx->host_scalar, u.scalar, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
               "max_dt readback");
        HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "max_dt sync");
        dt = std::min(*ctx->host_scalar, o->max_time_step) * cheby_step(u.n_cheby, i_cheby);
      }
      bool fixed = false;
      for (int i = 0; i < 2; ++i) {
        hx_kernel_options opts{dt, i, 0, o->use_filter};
        // face-flux folding (not with the admissibility fix, which reads the faces between the stages): stage 0 reads
        // the face states from the primary buffer and extrapolates into the second, stage 1 the other way round, so
        // that an update leaves the faces where the reference's sequence leaves them
        Stage_fold sf(mesh, !o->fix_admis && folds_faces(mesh), i);
        if ((s = hx_apply_state_bcs(mesh))) return s;
        if (o->use_ldg && !i) {
          s = hx_compute_navier_stokes(mesh, &opts, o->visc, o->cond);
        } else {
          if (i == 1 && u.fusable)
            mesh->fuse = Dt_fuse{u.block_min, 0, o->use_ldg ? 1 : 0, hx_basis_max_cfl(mesh->dm.rs) * u.msc,
                                 -2 / hx_basis_tables::eigenvalss[mesh->dm.rs - 2][1] * u.msd, u.ns};
          s = hx_compute_euler(mesh, &opts);
          mesh->fuse = Dt_fuse{};
        }
        if (s) return s;
        if (o->fix_admis) {
          int iters = 0;
          if ((s = hx_fix_admissibility(mesh, o, o->fix_admis_max_safety, &iters))) return s;
          st->fix_admis_iters += iters;
          fixed = iters || fixed;
        }
      }
      // a fix changed the state after the fused Max_dt read it, and ends this Chebyshev sequence (Solver.cpp:869)
      if (fixed) {
        pending = false;
        break;
      }
      pending = u.fusable;
      if (!device_dt) {
        st->time_step = dt;
        st->flow_time += dt;
      }
    }
  }
  if (!device_dt) ++st->iteration;
  return HX_OK;
}

// sets mesh->dt_dev for the duration of a device-form update
struct Device_dt
{
  hx_mesh* mesh;
  Device_dt(hx_mesh* m) : mesh{m} { mesh->dt_dev = &mesh->step->dt; }
  ~Device_dt() { mesh->dt_dev = nullptr; }
};

// the steady-state update (the fused minima pending when it starts, or local time stepping) captured into a graph
hx_status capture_update(hx_mesh* mesh, const hx_update_options* o, const Update_setup& u)
{
  hx_ctx* ctx = mesh->ctx;
  release_graph(mesh);
  HX_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  ctx->events_off = true;
  bool pending = u.fusable;
  hx_status s;
  {
    Device_dt d(mesh);
    s = update_once(mesh, o, u, nullptr, pending);
  }
  ctx->events_off = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (s) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  HX_TRY(ctx, e, "hipStreamEndCapture");
  const hipError_t ei = hipGraphInstantiate(&mesh->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) mesh->graph = nullptr;
  HX_TRY(ctx, ei, "hipGraphInstantiate");
  mesh->graph_opts = *o;
  mesh->graph_precision = mesh->precision;
  return HX_OK;
}

}  // namespace

extern "C" {

hx_status hx_update(hx_mesh* mesh, const hx_update_options* o, hx_update_status* st)
{
  if (!live(mesh) || !o || !st) return HX_ERR_INVALID;
  Update_setup u;
  if (hx_status s = update_setup(mesh, o, &u)) return s;
  bool pending = fused_pending(mesh, o, u);
  mesh->dt_cache.valid = false;
  if (hx_status s = update_once(mesh, o, u, st, pending)) return s;
  if (u.fusable && pending) remember_fused(mesh, o, u);
  return HX_OK;
}

hx_status hx_update_async(hx_mesh* mesh, const hx_update_options* o, int n_updates, const hx_update_status* from)
{
  if (!live(mesh) || !o || n_updates < 0) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  if (o->fix_admis)
    return fail(ctx, HX_ERR_UNSUPPORTED, "the admissibility fix reads its verdict on the host; use hx_update");
  if (mesh->n_cases) return fail(ctx, HX_ERR_INVALID, "the mesh holds a batch of cases; use hx_update_cases");
  Update_setup u;
  if (hx_status s = update_setup(mesh, o, &u)) return s;
  if (!mesh->step) {
    void* p = nullptr;
    HX_TRY(ctx, hipMalloc(&p, sizeof(Step_dev)), "hipMalloc(update status)");
    mesh->allocs.push_back(p);
    mesh->step = static_cast<Step_dev*>(p);
    hipLaunchKernelGGL(k_set_status, dim3(1), dim3(1), 0, ctx->stream, mesh->step, 0., 0., 0LL);
    HX_TRY(ctx, hipGetLastError(), "status");
  }
  if (from) {
    hipLaunchKernelGGL(k_set_status, dim3(1), dim3(1), 0, ctx->stream, mesh->step, from->time_step, from->flow_time,
                       (long long)from->iteration);
    HX_TRY(ctx, hipGetLastError(), "status");
  }
  bool pending = fused_pending(mesh, o, u);
  mesh->dt_cache.valid = false;
  for (int k = 0; k < n_updates; ++k) {
    if (u.fusable && !pending) {
      // the first update of a sequence computes Max_dt itself; every later one starts from the fused minima
      Device_dt d(mesh);
      if (hx_status s = update_once(mesh, o, u, nullptr, pending)) return s;
      continue;
    }
    if (!mesh->graph || mesh->graph_precision != mesh->precision || !same_options(mesh->graph_opts, *o))
      if (hx_status s = capture_update(mesh, o, u)) return s;
    Timer tm(ctx, "graph/update", mesh->dm.n_elem[0] + mesh->dm.n_elem[1]);
    HX_TRY(ctx, hipGraphLaunch(mesh->graph, ctx->stream), "hipGraphLaunch");
    ++mesh->gen;
    pending = u.fusable;
  }
  if (u.fusable && pending) remember_fused(mesh, o, u);
  return HX_OK;
}

hx_status hx_update_no_alloc(hx_mesh* mesh, const hx_update_options* o, hx_update_status* st)
{
  if (!live(mesh) || !o || !st) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  if (!ctx->scratch)
    return fail(ctx, HX_ERR_HIP, std::string("the scratch array has no device address: ") + hipGetErrorString(ctx->scratch_err));
  if (o->fix_admis)
    return fail(ctx, HX_ERR_UNSUPPORTED, "the admissibility fix allocates its flags; use hx_update");
  Update_setup u;
  if (hx_status s = update_setup(mesh, o, &u, true)) return s;
  bool pending = fused_pending(mesh, o, u);
  mesh->scratch_epoch = ++scratch_epochs[ctx->device];
  mesh->dt_cache.valid = false;
  // the kernels' reduction scalar is the array's, and no event is created for timing
  struct Scoped
  {
    hx_mesh* mesh;
    bool events;
    Scoped(hx_mesh* m) : mesh{m}, events{m->ctx->events_off}
    {
      mesh->scalar = mesh->ctx->scratch;
      mesh->ctx->events_off = true;
    }
    ~Scoped()
    {
      mesh->scalar = nullptr;
      mesh->ctx->events_off = events;
    }
  } scoped(mesh);
  if (hx_status s = update_once(mesh, o, u, st, pending)) return s;
  if (u.fusable && pending) remember_fused(mesh, o, u);
  return HX_OK;
}

hx_status hx_update_status_sync(hx_mesh* mesh, hx_update_status* st)
{
  if (!live(mesh) || !st) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  if (!mesh->step) return fail(ctx, HX_ERR_INVALID, "no asynchronous update has run on this mesh");
  static_assert(sizeof(Step_dev) <= 8 * sizeof(double), "the status fits the pinned readback buffer");
  HX_TRY(ctx, hipMemcpyAsync(ctx->host_scalar, mesh->step, sizeof(Step_dev), hipMemcpyDeviceToHost, ctx->stream),
         "status readback");
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "status sync");
  Step_dev d;
  std::memcpy(&d, ctx->host_scalar, sizeof(Step_dev));
  st->time_step = d.time_step;
  st->flow_time = d.flow_time;
  st->iteration = int(d.iteration);
  return HX_OK;
}

// Solver::compute_residual (Solver.cpp:899-913)
hx_status hx_compute_residual(hx_mesh* mesh, const hx_update_options* o)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_status s = hx_apply_state_bcs(mesh);
  if (s) return s;
  hx_kernel_options opts{1., 0, 1, o->use_filter};
  if (o->use_ldg) return hx_compute_navier_stokes(mesh, &opts, o->visc, o->cond);
  return hx_compute_euler(mesh, &opts);
}

// ---- batch of independent cases -------------------------------------------------------------------------------

hx_status hx_mesh_set_cases(hx_mesh* mesh, int n_cases, const int* case_cartesian, const int* case_deformed)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& m = mesh->dm;
  if (n_cases < 1) return fail(ctx, HX_ERR_INVALID, "n_cases must be positive");
  const int* ids[2] = {case_cartesian, case_deformed};
  for (int c = 0; c < 2; ++c) {
    if (m.n_elem[c] && !ids[c]) return fail(ctx, HX_ERR_INVALID, "missing case ids for a nonempty element class");
    for (int e = 0; e < m.n_elem[c]; ++e)
      if (ids[c][e] < 0 || ids[c][e] >= n_cases) return fail(ctx, HX_ERR_INVALID, "case id out of range");
  }
  // the cases must be independent: no face slot, connection or hanging face may join two of them
  HX_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  auto fetch = [&](const int* src, size_t n, std::vector<int>& dst) {
    dst.resize(n);
    return n ? hipMemcpy(dst.data(), src, n * sizeof(int), hipMemcpyDeviceToHost) : hipSuccess;
  };
  std::vector<int> slot_case(m.n_slot, -1), buf;
  for (int c = 0; c < 2; ++c) {
    HX_TRY(ctx, fetch(m.face_slot[c], size_t(m.n_elem[c]) * 2 * m.nd, buf), "download face slots");
    for (size_t k = 0; k < buf.size(); ++k) {
      int& sc = slot_case[buf[k]];
      const int id = ids[c][k / (2 * m.nd)];
      if (sc >= 0 && sc != id) return fail(ctx, HX_ERR_INVALID, "a face slot is shared by two cases");
      sc = id;
    }
  }
  for (int c = 0; c < 2; ++c) {
    HX_TRY(ctx, fetch(m.con_slot[c], size_t(m.n_con[c]) * 2, buf), "download connections");
    for (size_t k = 0; k < buf.size(); k += 2) {
      int &a = slot_case[buf[k]], &b = slot_case[buf[k + 1]];
      if (a >= 0 && b >= 0 && a != b) return fail(ctx, HX_ERR_INVALID, "a connection joins two cases");
      if (a < 0) a = b;  // ghost faces of boundary connections and hanging-face mortars take their partner's case
      if (b < 0) b = a;
    }
  }
  HX_TRY(ctx, fetch(m.ref_slot, size_t(m.n_ref) * 5, buf), "download refined faces");
  for (size_t k = 0; k < buf.size(); k += 5)
    for (int j = 1; j < 5; ++j)
      if (buf[k + j] >= 0 && slot_case[buf[k + j]] >= 0 && slot_case[buf[k + j]] != slot_case[buf[k]])
        return fail(ctx, HX_ERR_INVALID, "a hanging face joins two cases");
  auto alloc = [&](void** p, size_t bytes) {
    if (hipMalloc(p, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    mesh->allocs.push_back(*p);
    return true;
  };
  int* dev_ids[2] = {};
  void *cmin = nullptr, *cdt = nullptr;
  for (int c = 0; c < 2; ++c) {
    if (!alloc(reinterpret_cast<void**>(&dev_ids[c]), size_t(m.n_elem[c]) * sizeof(int)))
      return fail(ctx, HX_ERR_HIP, "hipMalloc(case ids)");
    if (m.n_elem[c])
      HX_TRY(ctx, hipMemcpy(dev_ids[c], ids[c], size_t(m.n_elem[c]) * sizeof(int), hipMemcpyHostToDevice), "upload case ids");
  }
  if (!alloc(&cmin, size_t(n_cases) * 8) || !alloc(&cdt, size_t(n_cases) * 8))
    return fail(ctx, HX_ERR_HIP, "hipMalloc(case time steps)");
  // cases listed in order within each class: their element ranges, for the streaming per-case reduction
  for (int c = 0; c < 2; ++c) {
    mesh->case_start[c] = nullptr;
    bool sorted = true;
    for (int e = 1; e < m.n_elem[c]; ++e) sorted = sorted && ids[c][e] >= ids[c][e - 1];
    if (!sorted || !m.n_elem[c]) continue;
    std::vector<int> start(n_cases + 1, 0);
    for (int e = 0; e < m.n_elem[c]; ++e) ++start[ids[c][e] + 1];
    int biggest = 0;
    for (int k = 0; k < n_cases; ++k) {
      biggest = std::max(biggest, start[k + 1]);
      start[k + 1] += start[k];
    }
    int* d = nullptr;
    if (!alloc(reinterpret_cast<void**>(&d), start.size() * sizeof(int))) return fail(ctx, HX_ERR_HIP, "hipMalloc(case ranges)");
    HX_TRY(ctx, hipMemcpy(d, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice), "upload case ranges");
    mesh->case_start[c] = d;
    mesh->case_split[c] = int(std::max<long long>(1, (long long)biggest * m.nq / 8192));
  }
  mesh->n_cases = n_cases;
  mesh->case_id[0] = dev_ids[0];
  mesh->case_id[1] = dev_ids[1];
  mesh->case_min = static_cast<unsigned long long*>(cmin);
  mesh->case_dt = static_cast<double*>(cdt);
  return HX_OK;
}

// Solver::update (Solver.cpp:834-886) of every case of the mesh at once. Each case takes its own global time step:
// the Max_dt minimum is reduced per case, each point's time step scale is set to its case's time step, and the stage
// kernels run with dt = 1, so that their coefficient times the time step scale is the same product as a single
// case's dt times a scale of 1.
}  // extern "C"

namespace {

// One time step of a batch of cases (hx_update_cases), split so that a group of batch meshes (hx_group) can run the
// per-mesh parts around its fused stage launches: Cases_step holds what one call needs; cases_setup prepares the mesh,
// cases_dt forms every case's time step on the device, cases_stage runs one stage on the mesh alone, cases_status
// advances the host status from the time steps read back, cases_end records the fused Max_dt for the next call.
struct Cases_step
{
  Launch l;
  bool visc;
  const char* name;
  double cc, cd;       // the constants of Max_dt with safety_conv = max_safety / max Chebyshev step, safety_diff = max_safety
  double n_cheby;
  bool fusable;        // every class lists its cases in order: stage 1's Local computes the next step's 1/scale per point
  bool pending;        // the previous call's per-point 1/scale is still that of the current state
  long long gen_at_entry;
};

hx_status cases_setup(hx_mesh* mesh, const hx_update_options* o, Cases_step& c)
{
  hx_ctx* ctx = mesh->ctx;
  c.gen_at_entry = mesh->gen;  // (make_launch below advances it)
  hx_pde p{};
  p.kind = o->use_ldg ? HX_PDE_NAVIER_STOKES : HX_PDE_EULER;
  p.visc = o->visc;
  p.cond = o->cond;
  c.l = make_launch(mesh);
  if (hx_status s = pde_check(mesh, &p, &c.visc, &c.l)) return s;
  const int rs = mesh->dm.rs;
  c.name = pde_name(o->use_ldg != 0);
  c.n_cheby = o->n_cheby_flow;
  const double max_cheby = cheby_step(c.n_cheby, c.n_cheby - 1);
  c.cc = hx_basis_max_cfl(rs) * (o->max_safety / max_cheby);
  c.cd = -2 / hx_basis_tables::eigenvalss[rs - 2][1] * o->max_safety;
  // with every class's cases in order, stage 1's Local also computes the next step's per-point 1/scale (the state it
  // writes is the state the next Max_dt reads); the next call uses it if nothing changed the mesh in between and the
  // time-step settings are the same
  c.fusable = true;
  for (int k = 0; k < 2; ++k) c.fusable = c.fusable && (!mesh->dm.n_elem[k] || mesh->case_start[k]);
  for (int k = 0; k < 2 && c.fusable; ++k)
    if (mesh->dm.n_elem[k] && !mesh->case_scale[k]) {
      void* q = nullptr;
      HX_TRY(ctx, hipMalloc(&q, size_t(mesh->dm.n_elem[k]) * mesh->dm.nq * sizeof(double)), "hipMalloc(case scale)");
      mesh->allocs.push_back(q);
      mesh->case_scale[k] = static_cast<double*>(q);
    }
  auto& cache = mesh->case_cache;
  c.pending = c.fusable && cache.valid && cache.gen == c.gen_at_entry && cache.cc == c.cc && cache.cd == c.cd &&
              cache.use_ldg == o->use_ldg && same_transport(cache.visc, o->visc) && same_transport(cache.cond, o->cond);
  cache.valid = false;
  return HX_OK;
}

// every case's time step min(Max_dt, max_time_step) * chebyshev_step into mesh->case_dt, and into the time step scale
hx_status cases_dt(hx_mesh* mesh, const hx_update_options* o, const Cases_step& c, int i_cheby)
{
  hx_ctx* ctx = mesh->ctx;
  const int nc = mesh->n_cases;
  hipLaunchKernelGGL(k_fill_u64, dim3((nc + 255) / 256), dim3(256), 0, c.l.stream, mesh->case_min,
                     (unsigned long long)__builtin_bit_cast(long long, DBL_MAX), nc);
  HX_TRY(ctx, hipGetLastError(), "fill");
  for (int k = 0; k < 2; ++k) {
    if (!mesh->dm.n_elem[k]) continue;
    Timer tm(ctx, std::string(cls_name(k)) + "/compute time step/" + c.name, mesh->dm.n_elem[k]);
    if (mesh->case_start[k]) {
      // each point's 1/scale into the time step scale (Max_dt in local mode), or the fused values of the previous
      // stage 1, then a streaming minimum per case
      if (!c.pending) HX_TRY(ctx, mesh->ks->max_dt(c.l, k, c.visc, 1, c.cc, c.cd), "max_dt");
      hipLaunchKernelGGL(k_case_min_tss, dim3(nc * mesh->case_split[k]), dim3(256), 0, c.l.stream,
                         c.pending ? mesh->case_scale[k] : mesh->dm.tss[k], mesh->case_start[k], mesh->dm.nq,
                         mesh->case_split[k], mesh->case_min);
      HX_TRY(ctx, hipGetLastError(), "case minimum");
    } else {
      HX_TRY(ctx, mesh->ks->max_dt_cases(c.l, k, c.visc, c.cc, c.cd, mesh->case_id[k], mesh->case_min), "max_dt_cases");
    }
  }
  hipLaunchKernelGGL(k_case_dt, dim3((nc + 255) / 256), dim3(256), 0, c.l.stream, mesh->case_min, mesh->case_dt, nc,
                     o->max_time_step, cheby_step(c.n_cheby, i_cheby));
  HX_TRY(ctx, hipGetLastError(), "case dt");
  for (int k = 0; k < 2; ++k) {
    const size_t n = size_t(mesh->dm.n_elem[k]) * mesh->dm.nq;
    if (!n) continue;
    hipLaunchKernelGGL(k_case_tss, dim3(unsigned(std::min<size_t>((n + 255) / 256, 4096))), dim3(256), 0, c.l.stream,
                       mesh->dm.tss[k], mesh->case_id[k], mesh->case_dt, mesh->dm.nq, n);
    HX_TRY(ctx, hipGetLastError(), "case tss");
  }
  return HX_OK;
}

// the stage-1 fusion of the Euler Local in a batch step: each point's next-step 1/scale into case_scale
Dt_fuse cases_fuse(const hx_mesh* mesh, const hx_update_options* o, const Cases_step& c)
{
  Dt_fuse f{};
  f.visc = o->use_ldg ? 1 : 0;
  f.max_cfl_c = c.cc;
  f.max_cfl_d = c.cd;
  f.ns = c.l.ns;
  f.point_scale[0] = mesh->case_scale[0];
  f.point_scale[1] = mesh->case_scale[1];
  return f;
}

// stage i of the mesh alone, each case's time step already in the time step scale (dt = 1), with face-flux folding
hx_status cases_stage(hx_mesh* mesh, const hx_update_options* o, const Cases_step& c, int i)
{
  hx_kernel_options opts{1., i, 0, o->use_filter};
  Stage_fold sf(mesh, true, i);
  hx_status s;
  if ((s = hx_apply_state_bcs(mesh))) return s;
  if (o->use_ldg && !i) return hx_compute_navier_stokes(mesh, &opts, o->visc, o->cond);
  if (i == 1 && c.fusable) mesh->fuse = cases_fuse(mesh, o, c);
  s = hx_compute_euler(mesh, &opts);
  mesh->fuse = Dt_fuse{};
  return s;
}

void cases_status(const hx_mesh* mesh, const double* dts, hx_update_status* st)
{
  for (int k = 0; k < mesh->n_cases; ++k) {
    st[k].time_step = dts[k];
    st[k].flow_time += dts[k];
  }
}

void cases_end(hx_mesh* mesh, const hx_update_options* o, const Cases_step& c, bool pending, hx_update_status* st)
{
  for (int k = 0; k < mesh->n_cases; ++k) ++st[k].iteration;
  if (c.fusable && pending) {
    auto& cache = mesh->case_cache;
    cache.valid = true;
    cache.gen = mesh->gen;
    cache.cc = c.cc;
    cache.cd = c.cd;
    cache.use_ldg = o->use_ldg;
    cache.visc = o->visc;
    cache.cond = o->cond;
  }
}

}  // namespace

extern "C" {

hx_status hx_update_cases(hx_mesh* mesh, const hx_update_options* o, hx_update_status* st)
{
  if (!live(mesh) || !o || !st) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  const int nc = mesh->n_cases;
  if (nc < 1) return fail(ctx, HX_ERR_INVALID, "no cases registered (hx_mesh_set_cases)");
  if (o->fix_admis) return fail(ctx, HX_ERR_UNSUPPORTED, "fix_admissibility is not built for batches of cases");
  if (o->local_time) {
    // every case takes the same step, min(1, max_time_step) * chebyshev_step
    hx_update_status one{};
    hx_status s = hx_update(mesh, o, &one);
    if (s) return s;
    for (int k = 0; k < nc; ++k) {
      st[k].time_step = one.time_step;
      st[k].flow_time += one.flow_time;
      st[k].iteration += one.iteration;
    }
    return HX_OK;
  }
  Cases_step c;
  if (hx_status s = cases_setup(mesh, o, c)) return s;
  std::vector<double> dts(nc);
  bool pending = c.pending;
  for (int i_flow = 0; i_flow < o->flow_iters; ++i_flow) {
    for (int i_cheby = 0; i_cheby < c.n_cheby; ++i_cheby) {
      c.pending = pending;
      if (hx_status s = cases_dt(mesh, o, c, i_cheby)) return s;
      for (int i = 0; i < 2; ++i)
        if (hx_status s = cases_stage(mesh, o, c, i)) return s;
      pending = c.fusable;
      HX_TRY(ctx, hipMemcpyAsync(dts.data(), mesh->case_dt, nc * sizeof(double), hipMemcpyDeviceToHost, c.l.stream),
             "case dt readback");
      HX_TRY(ctx, hipStreamSynchronize(c.l.stream), "case dt sync");
      cases_status(mesh, dts.data(), st);
    }
  }
  cases_end(mesh, o, c, pending, st);
  return HX_OK;
}

// ---- groups of meshes: fused launches over meshes, classes and row sizes ----------------------------------------

}  // extern "C"

namespace {

// the plan of one fused kernel: a segment per (mesh, class) with work, tiles from the fused set
void build_plan(hx_group* g, int kind, hx::Fused_plan& p)
{
  p = hx::Fused_plan{};
  for (int k = 0; k < int(g->meshes.size()); ++k)
    for (int c = 0; c < (kind == hx::fused_state_bcs ? 1 : 2); ++c) {
      const Dev_mesh& m = g->meshes[k]->dm;
      const int t = g->fs->tiles(kind, m, c);
      if (!t) continue;
      p.first[p.n_seg] = p.n_tiles;
      p.member[p.n_seg] = (signed char)k;
      p.cls[p.n_seg] = (signed char)c;
      p.rs[p.n_seg] = (signed char)m.rs;
      ++p.n_seg;
      p.n_tiles += t;
    }
  p.first[p.n_seg] = p.n_tiles;
}

// a fused launch of plan `kind` with the group's launch shape (the Local's shape; the face kernels one tile per
// workgroup)
hx::Fused_launch group_launch(hx_group* g, int kind, int stage)
{
  hx::Fused_launch l{};
  l.stream = g->ctx->stream;
  l.mem = g->mem + (stage ? g->meshes.size() : 0);
  l.plan = g->plans + kind;
  l.n_tiles = g->hplan[kind].n_tiles;
  l.order = kind == hx::fused_local ? g->order : 2;
  const long long cap = (long long)g->per_cu * g->n_cu;
  l.grid = unsigned(l.order == 2 || cap <= 0 ? l.n_tiles : std::min<long long>(cap, l.n_tiles));
  l.chunk = l.grid ? (l.n_tiles + int(l.grid) - 1) / int(l.grid) : 0;
  if (l.order == 1 && l.chunk) l.grid = unsigned((l.n_tiles + l.chunk - 1) / l.chunk);
  l.host_ops = g->hops;
  return l;
}

// the member tables: every mesh's device description and operator tables, and for stage 1 its Local's fused time step
hx_status upload_members(hx_group* g, const std::vector<Dt_fuse>& fuse)
{
  const int n = int(g->meshes.size());
  std::vector<hx::Fused_member> h(2 * n);
  for (int k = 0; k < n; ++k) {
    hx_mesh* m = g->meshes[k];
    // face-flux folding: stage 0 reads the face states from the primary buffer and extrapolates into the second, stage 1
    // the other way round
    h[k] = hx::Fused_member{m->dm, g->ctx->eops[m->dm.rs], g->ctx->ops[m->dm.rs], Dt_fuse{}};
    h[k].m.alt_rd = 0;
    h[k].m.alt_wr = 1;
    h[n + k] = h[k];
    h[n + k].m.alt_rd = 1;
    h[n + k].m.alt_wr = 0;
    h[n + k].fuse = fuse[k];
  }
  if (g->hmem.size() == h.size() && !std::memcmp(g->hmem.data(), h.data(), h.size() * sizeof(h[0]))) return HX_OK;
  HX_TRY(g->ctx, hipMemcpyAsync(g->mem, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, g->ctx->stream),
         "upload group members");
  HX_TRY(g->ctx, hipStreamSynchronize(g->ctx->stream), "upload group members");
  g->hmem = h;
  return HX_OK;
}

// one fused Euler stage of every mesh: state boundary conditions, Neighbor, (hanging faces per mesh) Local, prolong
hx_status group_stage(hx_group* g, const hx_update_options* o, int i)
{
  hx_ctx* ctx = g->ctx;
  const long long n_elem = [&] {
    long long s = 0;
    for (hx_mesh* m : g->meshes) s += (long long)m->dm.n_elem[0] + m->dm.n_elem[1];
    return s;
  }();
  hx_pde p{};
  p.kind = HX_PDE_EULER;
  for (hx_mesh* m : g->meshes) ++m->gen;
  // the stage's face buffers for the per-mesh hanging-face launches, reset when the stage ends
  struct Group_fold
  {
    const std::vector<hx_mesh*>& ms;
    Group_fold(const std::vector<hx_mesh*>& m, int stage) : ms{m}
    {
      for (hx_mesh* x : ms) {
        x->fold = 1;
        x->alt_rd = stage;
        x->alt_wr = 1 - stage;
      }
    }
    ~Group_fold()
    {
      for (hx_mesh* x : ms) x->fold = x->alt_rd = x->alt_wr = 0;
    }
  } folds(g->meshes, i);
  {
    Timer tm(ctx, "fused/boundary conditions/state", g->hplan[hx::fused_state_bcs].n_tiles);
    HX_TRY(ctx, g->fs->state_bcs(group_launch(g, hx::fused_state_bcs, i)), "fused state BCs");
  }
  {
    Timer tm(ctx, "fused/neighbor/euler", g->hplan[hx::fused_neighbor].n_tiles);
    HX_TRY(ctx, g->fs->neighbor(group_launch(g, hx::fused_neighbor, i)), "fused neighbor");
  }
  for (hx_mesh* m : g->meshes)
    if (m->dm.n_ref)
      if (hx_status s = hx_restrict_refined(m, &p, 1, 0)) return s;
  {
    Timer tm(ctx, "fused/local/euler", n_elem);
    const double coef = i ? hx_basis_step_ratio(0) : 1.;
    HX_TRY(ctx, g->fs->local_euler(group_launch(g, hx::fused_local, i), g->meshes[0]->precision, g->with_def, coef,
                                   nullptr, i, 0, o->use_filter),
           "fused local");
  }
  for (hx_mesh* m : g->meshes)
    if (m->dm.n_ref)
      if (hx_status s = hx_prolong_refined(m, &p, 0, 0)) return s;
  return HX_OK;
}

}  // namespace

extern "C" {

hx_status hx_group_create(hx_ctx* ctx, hx_mesh* const* meshes, int n_meshes, hx_group** out)
{
  if (!ctx || !meshes || !out || n_meshes < 1) return HX_ERR_INVALID;
  *out = nullptr;
  if (n_meshes > hx::fused_max_seg / 2) return fail(ctx, HX_ERR_INVALID, "a group holds at most 8 meshes");
  for (int k = 0; k < n_meshes; ++k) {
    if (!live(meshes[k]) || meshes[k]->ctx != ctx) return fail(ctx, HX_ERR_INVALID, "group meshes must be live meshes of the context");
    if (meshes[k]->n_cases < 1) return fail(ctx, HX_ERR_INVALID, "every mesh of a group needs its cases registered (hx_mesh_set_cases)");
    for (int j = 0; j < k; ++j)
      if (meshes[j] == meshes[k]) return fail(ctx, HX_ERR_INVALID, "a mesh appears twice in the group");
    if (meshes[k]->dm.nd != meshes[0]->dm.nd) return fail(ctx, HX_ERR_INVALID, "group meshes must share n_dim");
  }
  auto* g = new hx_group;
  g->ctx = ctx;
  g->meshes.assign(meshes, meshes + n_meshes);
  bool fusable = meshes[0]->dm.nd == 3;
  for (hx_mesh* m : g->meshes) {
    g->n_cases += m->n_cases;
    g->with_def = g->with_def || m->dm.n_elem[HX_DEFORMED] > 0;
    fusable = fusable && hx::fused_set_nd3()->supports(m->dm.rs) && !ensure_ops(ctx, m->dm.rs);
    g->hops[m->dm.rs] = ctx->host_ops[m->dm.rs].data();
  }
  auto bail = [&](hx_status s) {
    release_group(g);
    delete g;
    return s;
  };
  if (hipHostMalloc(reinterpret_cast<void**>(&g->dts), std::max(1, g->n_cases) * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(ctx, HX_ERR_HIP, "hipHostMalloc(group time steps)"));
  if (fusable) {
    g->fs = hx::fused_set_nd3();
    if (hipMalloc(reinterpret_cast<void**>(&g->mem), 2 * n_meshes * sizeof(hx::Fused_member)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&g->plans), 3 * sizeof(hx::Fused_plan)) != hipSuccess)
      return bail(fail(ctx, HX_ERR_HIP, "hipMalloc(group tables)"));
    for (int kind = 0; kind < 3; ++kind) build_plan(g, kind, g->hplan[kind]);
    if (hipMemcpy(g->plans, g->hplan, sizeof(g->hplan), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(ctx, HX_ERR_HIP, "upload group plans"));
    if (hx_status s = upload_members(g, std::vector<Dt_fuse>(n_meshes))) return bail(s);
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
      return bail(fail(ctx, HX_ERR_HIP, "hipGetDeviceProperties"));
    g->n_cu = prop.multiProcessorCount;
  }
  for (hx_mesh* m : g->meshes) ++m->n_groups;
  ctx->groups.push_back(g);
  *out = g;
  return HX_OK;
}

hx_status hx_group_destroy(hx_group* g)
{
  if (!g) return HX_OK;
  if (hx_ctx* ctx = g->ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->groups.erase(std::remove(ctx->groups.begin(), ctx->groups.end(), g), ctx->groups.end());
    release_group(g);
  }
  delete g;
  return HX_OK;
}

int hx_group_fused(const hx_group* g) { return g && g->ctx && g->fs ? 1 : 0; }

hx_status hx_group_set_shape(hx_group* g, int order, int workgroups_per_cu)
{
  if (!g || !g->ctx) return HX_ERR_INVALID;
  if (order < 0 || order > 2 || (order != 2 && workgroups_per_cu < 1))
    return fail(g->ctx, HX_ERR_INVALID, "launch shape: order 0 or 1 with workgroups_per_cu >= 1, or order 2");
  g->order = order;
  g->per_cu = order == 2 ? 0 : workgroups_per_cu;
  return HX_OK;
}

hx_status hx_group_update(hx_group* g, const hx_update_options* o, hx_update_status* st)
{
  if (!g || !g->ctx || !o || !st) return HX_ERR_INVALID;
  hx_ctx* ctx = g->ctx;
  const int n = int(g->meshes.size());
  for (hx_mesh* m : g->meshes)
    if (m->precision != g->meshes[0]->precision)
      return fail(ctx, HX_ERR_INVALID, "group meshes must share the arithmetic form");
  // the fused launches cover Euler stages with a global time step; anything else steps the meshes one at a time
  if (!g->fs || o->use_ldg || o->local_time || o->fix_admis) {
    for (int k = 0, off = 0; k < n; off += g->meshes[k]->n_cases, ++k)
      if (hx_status s = hx_update_cases(g->meshes[k], o, st + off)) return s;
    return HX_OK;
  }
  std::vector<Cases_step> c(n);
  std::vector<Dt_fuse> fuse(n);
  std::vector<char> pending(n);
  for (int k = 0; k < n; ++k) {
    if (hx_status s = cases_setup(g->meshes[k], o, c[k])) return s;
    fuse[k] = c[k].fusable ? cases_fuse(g->meshes[k], o, c[k]) : Dt_fuse{};
    pending[k] = c[k].pending;
  }
  if (hx_status s = upload_members(g, fuse)) return s;
  for (int i_flow = 0; i_flow < o->flow_iters; ++i_flow) {
    for (int i_cheby = 0; i_cheby < c[0].n_cheby; ++i_cheby) {
      for (int k = 0; k < n; ++k) {
        c[k].pending = pending[k];
        if (hx_status s = cases_dt(g->meshes[k], o, c[k], i_cheby)) return s;
      }
      for (int i = 0; i < 2; ++i)
        if (hx_status s = group_stage(g, o, i)) return s;
      for (int k = 0, off = 0; k < n; off += g->meshes[k]->n_cases, ++k) {
        pending[k] = c[k].fusable;
        HX_TRY(ctx, hipMemcpyAsync(g->dts + off, g->meshes[k]->case_dt, g->meshes[k]->n_cases * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream), "case dt readback");
      }
      HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "case dt sync");
      for (int k = 0, off = 0; k < n; off += g->meshes[k]->n_cases, ++k) cases_status(g->meshes[k], g->dts + off, st + off);
    }
  }
  for (int k = 0, off = 0; k < n; off += g->meshes[k]->n_cases, ++k) cases_end(g->meshes[k], o, c[k], pending[k], st + off);
  return HX_OK;
}

hx_status hx_group_autotune(hx_group* g, int use_filter, int reps, hx_launch_shape* table, int capacity, int* n_table)
{
  if (!g || !g->ctx || reps < 1) return HX_ERR_INVALID;
  hx_ctx* ctx = g->ctx;
  if (!g->fs) return fail(ctx, HX_ERR_UNSUPPORTED, "the group has no fused launches to tune");
  // candidates: one tile per workgroup, and persistent grids of 2 to 16 workgroups per compute unit walking the tiles
  // strided or in contiguous chunks
  std::vector<hx_launch_shape> cand{{2, 0, 0.}};
  for (int order = 0; order < 2; ++order)
    for (int per_cu : {2, 4, 8, 16}) cand.push_back({order, per_cu, 0.});
  hipEvent_t a = nullptr, b = nullptr;
  HX_TRY(ctx, hipEventCreate(&a), "hipEventCreate");
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(ctx, HX_ERR_HIP, "hipEventCreate");
  }
  const int keep_order = g->order, keep_per_cu = g->per_cu;
  hx_status status = HX_OK;
  for (auto& cs : cand) {
    g->order = cs.order;
    g->per_cu = cs.workgroups_per_cu;
    // the Local in residual mode (stage 0 tables, no time-step fusion): the state is left as it is, the residual cache
    // is overwritten and the faces get the extrapolation of the unchanged state
    auto run = [&] {
      return g->fs->local_euler(group_launch(g, hx::fused_local, 0), g->meshes[0]->precision, g->with_def, 1., nullptr,
                                0, 1, use_filter);
    };
    hipError_t e = run();  // warm-up
    float best = 0.f;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      float ms = 0.f;
      e = hipEventRecord(a, ctx->stream);
      if (e == hipSuccess) e = run();
      if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
      if (e == hipSuccess && (r == 0 || ms < best)) best = ms;
    }
    if (e != hipSuccess) {
      status = fail(ctx, HX_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
      break;
    }
    cs.ms = best;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  for (hx_mesh* m : g->meshes) ++m->gen;
  if (status) {
    g->order = keep_order;
    g->per_cu = keep_per_cu;
    return status;
  }
  const auto best = std::min_element(cand.begin(), cand.end(), [](const auto& x, const auto& y) { return x.ms < y.ms; });
  g->order = best->order;
  g->per_cu = best->workgroups_per_cu;
  if (n_table) *n_table = int(cand.size());
  if (table)
    for (int k = 0; k < std::min(capacity, int(cand.size())); ++k) table[k] = cand[k];
  return HX_OK;
}

// ---- artificial-viscosity generator ---------------------------------------------------------------------------

hx_status hx_apply_advection_bcs(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Timer tm(mesh->ctx, "boundary conditions/advection", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->av_bcs(make_launch(mesh), 0), "advection bcs");
  return HX_OK;
}

hx_status hx_apply_avc_diff_bcs(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Timer tm(mesh->ctx, "boundary conditions/av diffusion", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->av_bcs(make_launch(mesh), 1), "avc diff bcs");
  return HX_OK;
}

hx_status hx_apply_avc_diff_flux_bcs(hx_mesh* mesh)
{
  if (!live(mesh)) return HX_ERR_INVALID;
  Timer tm(mesh->ctx, "boundary conditions/av diffusion flux", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->av_bcs(make_launch(mesh), 2), "avc diff flux bcs");
  return HX_OK;
}

hx_status hx_compute_advection(hx_mesh* mesh, const hx_kernel_options* o, double advect_length)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  return compute_advection_impl(mesh, o, advect_length, 0);
}

hx_status hx_compute_smooth_av(hx_mesh* mesh, const hx_kernel_options* o, double diff_time, double cheby_step)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_SMOOTH_AV;
  p.diff_time = diff_time;
  p.cheby_step = cheby_step;
  if (o->i_stage)
    return fail(mesh->ctx, HX_ERR_INVALID, "two-stage stabilization is not applicable to diffusion equations");
  hx_status s;
  if ((s = hx_neighbor(mesh, HX_CARTESIAN, &p, 0))) return s;
  if ((s = hx_neighbor(mesh, HX_DEFORMED, &p, 0))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 0))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 0, 1))) return s;
  if ((s = hx_local(mesh, HX_CARTESIAN, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_local(mesh, HX_DEFORMED, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_prolong_refined(mesh, &p, 1, 1))) return s;
  if ((s = hx_apply_avc_diff_flux_bcs(mesh))) return s;
  if ((s = hx_neighbor_reconcile(mesh, HX_CARTESIAN, &p))) return s;
  if ((s = hx_neighbor_reconcile(mesh, HX_DEFORMED, &p))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 1))) return s;
  if ((s = hx_reconcile_ldg_flux(mesh, HX_CARTESIAN, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_reconcile_ldg_flux(mesh, HX_DEFORMED, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  return hx_prolong_refined(mesh, &p, 0, 0);
}

hx_status hx_compute_write_face_advection(hx_mesh* mesh, double advect_length)
{
  hx_pde p{};
  p.kind = HX_PDE_ADVECTION;
  p.advect_length = advect_length;
  hx_status s;
  if ((s = hx_write_face(mesh, HX_CARTESIAN, &p))) return s;
  return hx_write_face(mesh, HX_DEFORMED, &p);
}

hx_status hx_compute_write_face_smooth_av(hx_mesh* mesh)
{
  hx_pde p{};
  p.kind = HX_PDE_SMOOTH_AV;
  p.diff_time = p.cheby_step = 1.;  // Write_face(basis, 1., 1.) (kernels_convective.cpp:53-56)
  hx_status s;
  if ((s = hx_write_face(mesh, HX_CARTESIAN, &p))) return s;
  return hx_write_face(mesh, HX_DEFORMED, &p);
}

hx_status hx_compute_prolong_advection(hx_mesh* mesh, double advect_length)
{
  hx_pde p{};
  p.kind = HX_PDE_ADVECTION;
  p.advect_length = advect_length;
  return hx_prolong_refined(mesh, &p, 0, 0);
}

hx_status hx_max_dt_advection(hx_mesh* mesh, const hx_kernel_options* o, double msc, double msd, int local_time,
                              double advect_length, double* dt_out)
{
  if (!live(mesh) || !dt_out) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_ADVECTION;
  p.advect_length = advect_length;
  double a, b;
  hx_status s;
  const int uf = o ? o->use_filter : 0;
  if ((s = hx_max_dt(mesh, HX_CARTESIAN, &p, local_time, uf, msc, msd, &a))) return s;
  if ((s = hx_max_dt(mesh, HX_DEFORMED, &p, local_time, uf, msc, msd, &b))) return s;
  *dt_out = std::min(a, b);
  return HX_OK;
}

hx_status hx_max_dt_smooth_av(hx_mesh* mesh, const hx_kernel_options* o, double msc, double msd, int local_time,
                              double* dt_out)
{
  if (!live(mesh) || !dt_out) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_SMOOTH_AV;
  p.diff_time = p.cheby_step = 1.;  // Max_dt<Smooth_art_visc>(..., 1., 1.) (kernels_max_dt.cpp:15)
  double a, b;
  hx_status s;
  const int uf = o ? o->use_filter : 0;
  if ((s = hx_max_dt(mesh, HX_CARTESIAN, &p, local_time, uf, msc, msd, &a))) return s;
  if ((s = hx_max_dt(mesh, HX_DEFORMED, &p, local_time, uf, msc, msd, &b))) return s;
  *dt_out = std::min(a, b);
  return HX_OK;
}

// The fix's boundary faces: mode 1 ghost state = inside (Solver.cpp:1066-1071), mode 2 ghost LDG = -inside
// (Solver::apply_fta_flux_bcs, Solver.cpp:103-115); the same copies as the AV diffusion hooks
static hx_status fta_bcs(hx_mesh* mesh, int mode)
{
  Timer tm(mesh->ctx, "fix admis./boundary conditions", mesh->dm.n_bc_con);
  HX_TRY(mesh->ctx, mesh->ks->av_bcs(make_launch(mesh), mode), "fix admissibility bcs");
  return HX_OK;
}

// ---- Solver::calc_jacobian (Solver.cpp:273-381) ---------------------------------------------------------------------

}  // extern "C"

namespace {

// every vertex index a vertex-sharing kernel follows: vertex ids and the hanging-vertex matchers
hx_status check_vertex_data(hx_mesh* mesh, const hx_vertex_data* v)
{
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  const int nd = dm.nd, nvert = 1 << nd;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1];
  if (v->n_vertex < 0 || (n_tot && v->n_vertex == 0)) return fail(ctx, HX_ERR_INVALID, "vertex count missing");
  for (int c = 0; c < 2; ++c) {
    if (dm.n_elem[c] && !v->vertex_id[c]) return fail(ctx, HX_ERR_INVALID, "vertex ids missing");
    for (long long k = 0; k < (long long)dm.n_elem[c] * nvert; ++k)
      if (v->vertex_id[c][k] < 0 || v->vertex_id[c][k] >= v->n_vertex)
        return fail(ctx, HX_ERR_INVALID, "vertex id out of range");
  }
  if (v->n_hanging < 0 || (v->n_hanging && !v->hanging)) return fail(ctx, HX_ERR_INVALID, "hanging vertices missing");
  if (v->n_hanging && nd < 2) return fail(ctx, HX_ERR_INVALID, "hanging vertices need n_dim >= 2");
  for (int i = 0; i < v->n_hanging; ++i) {
    const int* h = v->hanging + 12 * i;
    if (h[0] < 0 || h[0] >= nd || (h[1] != 0 && h[1] != 1)) return fail(ctx, HX_ERR_INVALID, "bad hanging-vertex face");
    int n_fine = 1 << (nd - 1);
    for (int d = 0; d < nd - 1; ++d)
      if (h[2 + d]) n_fine /= 2;
    for (int f = 0; f < n_fine; ++f) {
      const int c = h[4 + 2 * f], e = h[5 + 2 * f];
      if (c < 0 || c > 1 || e < 0 || e >= dm.n_elem[c]) return fail(ctx, HX_ERR_INVALID, "hanging-vertex element out of range");
    }
  }
  return HX_OK;
}

// device buffer `*p` of `count` elements, (re)allocated when `*p` is null or `realloc`
template <class T>
hipError_t vc_buffer(T** p, size_t count, bool realloc)
{
  if (*p && !realloc) return hipSuccess;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  return count ? hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) : hipSuccess;
}

// makes the mesh's vertex cache hold the topology `v`: a topology equal to the cached one (same vertex count, ids and
// hanging vertices) is taken as it is; another is validated, copied and uploaded. Also allocates the element work
// arrays (`vert`, `elem`) on first use.
hx_status sync_vertices(hx_mesh* mesh, const hx_vertex_data* v)
{
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  auto& vc = mesh->vc;
  const int nvert = 1 << dm.nd;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1], n = n_tot * nvert;
  HX_TRY(ctx, vc_buffer(&vc.vert, size_t(n), false), "hipMalloc(vertex values)");
  HX_TRY(ctx, vc_buffer(&vc.elem, size_t(n_tot), false), "hipMalloc(element values)");
  bool same = vc.n_vertex >= 0 && v->n_vertex == vc.n_vertex && v->n_hanging >= 0 &&
              (long long)vc.hanging.size() == 12LL * v->n_hanging && (long long)vc.ids.size() == n;
  for (int c = 0, off = 0; same && c < 2; off += dm.n_elem[c] * nvert, ++c)
    if (dm.n_elem[c])
      same = v->vertex_id[c] && !std::memcmp(vc.ids.data() + off, v->vertex_id[c], size_t(dm.n_elem[c]) * nvert * sizeof(int));
  if (same && v->n_hanging)
    same = v->hanging && !std::memcmp(vc.hanging.data(), v->hanging, vc.hanging.size() * sizeof(int));
  if (same) return HX_OK;
  if (hx_status s = check_vertex_data(mesh, v)) return s;
  // the stream's earlier work may still read the buffers replaced below; from here on the cache is invalid until the
  // new topology is on the device
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "vertex upload sync");
  vc.n_vertex = -1;
  vc.ids.resize(size_t(n));
  for (int c = 0, off = 0; c < 2; off += dm.n_elem[c] * nvert, ++c)
    if (dm.n_elem[c]) std::memcpy(vc.ids.data() + off, v->vertex_id[c], size_t(dm.n_elem[c]) * nvert * sizeof(int));
  vc.hanging.assign(v->hanging, v->hanging + 12LL * v->n_hanging);
  HX_TRY(ctx, vc_buffer(&vc.vid, size_t(n), false), "hipMalloc(vertex ids)");
  if (vc.cap_vertex < v->n_vertex) {
    vc.cap_vertex = 0;
    HX_TRY(ctx, vc_buffer(&vc.shared, size_t(v->n_vertex), true), "hipMalloc(shared vertices)");
    vc.cap_vertex = v->n_vertex;
  }
  if (vc.cap_hanging < v->n_hanging) {
    vc.cap_hanging = 0;
    HX_TRY(ctx, vc_buffer(&vc.hang, size_t(12LL * v->n_hanging), true), "hipMalloc(hanging vertices)");
    HX_TRY(ctx, vc_buffer(&vc.vals, size_t(9LL * v->n_hanging), true), "hipMalloc(hanging values)");
    vc.cap_hanging = v->n_hanging;
  }
  if (n) HX_TRY(ctx, hipMemcpy(vc.vid, vc.ids.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice), "upload vertex ids");
  if (v->n_hanging)
    HX_TRY(ctx, hipMemcpy(vc.hang, vc.hanging.data(), vc.hanging.size() * sizeof(int), hipMemcpyHostToDevice),
           "upload hanging vertices");
  vc.n_vertex = v->n_vertex;
  return HX_OK;
}

// the per-element vertex copies `vert` [n_cartesian + n_deformed][2^nd] shared over every vertex (by minimum or by
// maximum), then the hanging vertices matched, over the topology in the mesh's vertex cache (sync_vertices)
hx_status share_vertices(hx_mesh* mesh, double* vert, bool maximum)
{
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  const auto& vc = mesh->vc;
  const int nvert = 1 << dm.nd;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1], n = n_tot * nvert;
  const int n_hanging = int(vc.hanging.size() / 12);
  if (n) {
    const double inf = std::numeric_limits<double>::infinity();
    hipLaunchKernelGGL(k_fill_n, dim3(unsigned((vc.n_vertex + 255) / 256)), dim3(256), 0, ctx->stream, vc.shared,
                       maximum ? -inf : inf, vc.n_vertex);
    if (maximum)
      hipLaunchKernelGGL(k_vertex_push_max, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, vc.vid, vert, n, vc.shared);
    else
      hipLaunchKernelGGL(k_vertex_push_min, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, vc.vid, vert, n, vc.shared);
    hipLaunchKernelGGL(k_vertex_fetch, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, vc.vid, vert, n, vc.shared);
    HX_TRY(ctx, hipGetLastError(), "vertex sharing");
  }
  Launch l = make_launch(mesh);
  HX_TRY(ctx, mesh->ks->hanging_match(l, vc.hang, n_hanging, vert, vc.vals, 0), "hanging vertices");
  HX_TRY(ctx, mesh->ks->hanging_match(l, vc.hang, n_hanging, vert, vc.vals, 1), "hanging vertices");
  return HX_OK;
}

// Stab_art_visc (stabilizing_art_visc.cpp:8-61) over both classes into a device array [n_tot][2^nd] (each element's
// value in all its vertex slots), and the element values copied to `uncert` when it is not null
hx_status stab_into(hx_mesh* mesh, double char_speed, double* vert, double* uncert)
{
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  const int nvert = 1 << dm.nd;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1];
  Launch l = make_launch(mesh);
  for (int cls = 0; cls < 2; ++cls) {
    Timer tm(ctx, std::string(cls_name(cls)) + "/stabilizing art visc/indicator", dm.n_elem[cls]);
    HX_TRY(ctx, mesh->ks->stab_indicator(l, cls, char_speed, vert), "stab_art_visc");
  }
  if (uncert && n_tot) {
    // each element's value is in all its vertex slots: the first slot of each, gathered on the device
    double* elem = mesh->vc.elem;
    hipLaunchKernelGGL(k_gather_stride, dim3(unsigned((n_tot + 255) / 256)), dim3(256), 0, ctx->stream, vert, nvert,
                       n_tot, elem);
    HX_TRY(ctx, hipGetLastError(), "element values");
    HX_TRY(ctx, hipMemcpyAsync(uncert, elem, size_t(n_tot) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
           "element values readback");
    HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "stab_art_visc sync");
  }
  return HX_OK;
}

}  // namespace

extern "C" {

hx_status hx_calc_jacobian(hx_mesh* mesh, const hx_vertex_data* v)
{
  if (!live(mesh) || !v) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  Dev_mesh& dm = mesh->dm;
  const int nd = dm.nd, rs = dm.rs, nvert = 1 << nd, nfq = dm.nfq;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1];
  if (hx_status s = sync_vertices(mesh, v)) return s;
  if (dm.n_elem[HX_DEFORMED] && !v->position) return fail(ctx, HX_ERR_INVALID, "vertex positions missing");
  ++mesh->gen;
  Scratch sc;
  double *pos = nullptr, *adj = nullptr, *vt = nullptr;
  const long long n_def = dm.n_elem[HX_DEFORMED];
  HX_TRY(ctx, sc.put(&pos, v->position, size_t(n_def * nvert * nd)), "upload vertex positions");
  if (v->node_adj) HX_TRY(ctx, sc.put(&adj, v->node_adj, size_t(n_def * 2 * nd * nfq)), "upload node adjustments");
  HX_TRY(ctx, sc.put(&vt, (const double*)nullptr, size_t(n_tot * nvert)), "hipMalloc(vertex scale)");
  Launch l = make_launch(mesh);
  {
    Timer tm(ctx, "compute jacobian/elements", n_tot);
    const int k = rs - 2;
    HX_TRY(ctx, mesh->ks->set_jacobian(l, pos, adj, vt, hx_basis_tables::nodes[k], hx_basis_tables::diff_mats[k],
                                       hx_basis_tables::boundarys[k]), "set_jacobian");
  }
  {
    Timer tm(ctx, "compute jacobian/normals", dm.n_con[HX_DEFORMED]);
    HX_TRY(ctx, mesh->ks->connection_normals(l), "connection normals");
  }
  {
    Timer tm(ctx, "compute jacobian/vertex time step scale", n_tot);
    if (hx_status s = share_vertices(mesh, vt, false)) return s;
  }
  for (int c = 0, off = 0; c < 2; off += dm.n_elem[c] * nvert, ++c)
    if (dm.n_elem[c])
      HX_TRY(ctx, hipMemcpyAsync(dm.vertex_tss[c], vt + off, size_t(dm.n_elem[c]) * nvert * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx->stream), "vertex scale copy");
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "calc_jacobian sync");
  return HX_OK;
}


hx_status hx_stab_art_visc(hx_mesh* mesh, double char_speed, double* uncert)
{
  if (!live(mesh) || !uncert) return HX_ERR_INVALID;
  auto& vc = mesh->vc;
  const long long n_tot = (long long)mesh->dm.n_elem[0] + mesh->dm.n_elem[1];
  HX_TRY(mesh->ctx, vc_buffer(&vc.vert, size_t(n_tot) << mesh->dm.nd, false), "hipMalloc(vertex values)");
  HX_TRY(mesh->ctx, vc_buffer(&vc.elem, size_t(n_tot), false), "hipMalloc(element values)");
  return stab_into(mesh, char_speed, vc.vert, uncert);
}

hx_status hx_set_art_visc_admis(hx_mesh* mesh, const hx_vertex_data* v, double char_speed, double* uncert)
{
  if (!live(mesh) || !v) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  for (int c = 0; c < 2; ++c)
    if (dm.n_elem[c] && !dm.av[c]) return fail(ctx, HX_ERR_INVALID, "set_art_visc_admis needs the artificial viscosity arrays");
  if (hx_status s = sync_vertices(mesh, v)) return s;
  const long long n_tot = (long long)dm.n_elem[0] + dm.n_elem[1];
  double* vert = mesh->vc.vert;
  if (hx_status s = stab_into(mesh, char_speed, vert, uncert)) return s;
  {
    Timer tm(ctx, "stabilizing art visc/vertices", n_tot);
    if (hx_status s = share_vertices(mesh, vert, true)) return s;
  }
  Launch l = make_launch(mesh);
  for (int cls = 0; cls < 2; ++cls) {
    Timer tm(ctx, std::string(cls_name(cls)) + "/stabilizing art visc/interpolate", dm.n_elem[cls]);
    HX_TRY(ctx, mesh->ks->vertex_interp_lap(l, cls, vert), "vertex interpolation");
  }
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "set_art_visc_admis sync");
  return HX_OK;
}

// compute_fix_therm_admis (kernels_diffusive.cpp:32): the COMPUTE_DIFFUSION sequence with pde::Fix_therm_admis; the
// flux_bc callback is the fix's own (Solver::apply_fta_flux_bcs, Solver.cpp:103-115)
hx_status hx_compute_fix_therm_admis(hx_mesh* mesh, const hx_kernel_options* o)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_FIX_ADMIS;
  if (o->i_stage)
    return fail(mesh->ctx, HX_ERR_INVALID, "two-stage stabilization is not applicable to diffusion equations");
  hx_status s;
  if ((s = hx_neighbor(mesh, HX_CARTESIAN, &p, 0))) return s;
  if ((s = hx_neighbor(mesh, HX_DEFORMED, &p, 0))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 0))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 0, 1))) return s;
  if ((s = hx_local(mesh, HX_CARTESIAN, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_local(mesh, HX_DEFORMED, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_prolong_refined(mesh, &p, 1, 1))) return s;
  if ((s = fta_bcs(mesh, 2))) return s;
  if ((s = hx_neighbor_reconcile(mesh, HX_CARTESIAN, &p))) return s;
  if ((s = hx_neighbor_reconcile(mesh, HX_DEFORMED, &p))) return s;
  if ((s = hx_restrict_refined(mesh, &p, 1, 1))) return s;
  if ((s = hx_reconcile_ldg_flux(mesh, HX_CARTESIAN, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  if ((s = hx_reconcile_ldg_flux(mesh, HX_DEFORMED, &p, o->dt, 0, o->compute_residual, o->use_filter))) return s;
  return hx_prolong_refined(mesh, &p, 0, 0);
}

hx_status hx_max_dt_fix_therm_admis(hx_mesh* mesh, const hx_kernel_options* o, double msc, double msd, int local_time,
                                    double* dt_out)
{
  if (!live(mesh) || !dt_out) return HX_ERR_INVALID;
  hx_pde p{};
  p.kind = HX_PDE_FIX_ADMIS;
  double a, b;
  hx_status s;
  const int uf = o ? o->use_filter : 0;
  if ((s = hx_max_dt(mesh, HX_CARTESIAN, &p, local_time, uf, msc, msd, &a))) return s;
  if ((s = hx_max_dt(mesh, HX_DEFORMED, &p, local_time, uf, msc, msd, &b))) return s;
  *dt_out = std::min(a, b);
  return HX_OK;
}

// Solver::is_admissible (Solver.cpp:921-959)
hx_status hx_is_admissible(hx_mesh* mesh, int* admissible)
{
  if (!live(mesh) || !admissible) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  // reads the fields only, so a fused Max_dt stays valid
  if (!mesh->admis_flags) {
    void* p = nullptr;
    HX_TRY(ctx, hipMalloc(&p, 2 * sizeof(int)), "hipMalloc(admissibility flags)");
    mesh->allocs.push_back(p);
    mesh->admis_flags = static_cast<int*>(p);
  }
  Launch l = make_launch(mesh, false);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, ctx->stream, reinterpret_cast<unsigned long long*>(mesh->admis_flags),
                     0ull, 1);
  HX_TRY(ctx, hipGetLastError(), "fill");
  {
    Timer tm(ctx, "check admis.", (long long)mesh->dm.n_elem[0] + mesh->dm.n_elem[1]);
    HX_TRY(ctx, mesh->ks->admissible(l, mesh->admis_flags), "is_admissible");
  }
  int h[2] = {0, 0};
  HX_TRY(ctx, hipMemcpyAsync(h, mesh->admis_flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "admissibility readback");
  HX_TRY(ctx, hipStreamSynchronize(ctx->stream), "admissibility sync");
  if (h[1]) return fail(ctx, HX_ERR_NOT_FINITE, "state is not finite");
  *admissible = !h[0];
  return HX_OK;
}

// Solver::fix_admissibility (Solver.cpp:961-1095). Each iteration of the reference writes a fix coefficient (the
// inadmissible elements' flags, shared by maximum over vertices twice, matched at hanging vertices and interpolated to
// the quadrature points) into the Laplacian artificial viscosity and swaps it with the bulk one around the fix's Max_dt,
// its two diffusion steps and the flow's Max_dt; afterwards the Laplacian coefficient holds the fix coefficient. The
// coefficient reaches no result: the fix's Max_dt and diffusion do not read artificial viscosity; the flow Max_dt's time
// step scale is overwritten by the next iteration's fix Max_dt; and the last iteration always starts from an
// admissible state (an inadmissible one restarts the count), so its coefficient is zero everywhere. Here every
// iteration therefore swaps with a zero coefficient, which leaves the same state, time step scale and artificial
// viscosity (Laplacian coefficient zero) as the reference.
hx_status hx_fix_admissibility(hx_mesh* mesh, const hx_update_options* o, double stability_ratio, int* n_iters_out)
{
  if (!live(mesh) || !o || !n_iters_out) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  Dev_mesh& dm = mesh->dm;
  hx_status s;
  int iter;
  int n_iters = std::numeric_limits<int>::max();
  for (iter = 0; iter < n_iters; ++iter) {
    if (iter >= 100000)
      return fail(ctx, HX_ERR_INVALID,
                  "failed to fix thermodynamic admissability in " + std::to_string(iter) + " iterations");
    int admissible = 0;
    if ((s = hx_is_admissible(mesh, &admissible))) return s;
    if (admissible) {
      if (iter) n_iters = std::min(n_iters, 2 * iter);
      else {
        ++iter;
        break;
      }
    } else {
      n_iters = std::numeric_limits<int>::max();
    }
    // every element has artificial-viscosity storage in the reference (zero unless set)
    for (int c = 0; c < 2; ++c)
      if (dm.n_elem[c] && !dm.av[c]) {
        void* p = nullptr;
        const size_t bytes = size_t(dm.n_elem[c]) * 2 * dm.nq * sizeof(double);
        HX_TRY(ctx, hipMalloc(&p, bytes), "hipMalloc(artificial viscosity)");
        mesh->allocs.push_back(p);
        HX_TRY(ctx, hipMemsetAsync(p, 0, bytes, ctx->stream), "hipMemsetAsync(artificial viscosity)");
        dm.av[c] = static_cast<double*>(p);
      }
    Launch l = make_launch(mesh);
    HX_TRY(ctx, mesh->ks->fix_av(l, 0), "fix_av");
    hx_kernel_options ko{0., 0, 0, 0};
    double dummy;
    if ((s = hx_max_dt_fix_therm_admis(mesh, &ko, stability_ratio, stability_ratio, 1, &dummy))) return s;
    // dt = 1: the two steps of a second-order stabilized update (Solver.cpp:1059-1064)
    const double linear = 1., quadratic = 1. / 8 / 0.9;
    double step[2];
    step[1] = (linear + std::sqrt(linear * linear - 4 * quadratic)) / 2.;
    step[0] = quadratic / step[1];
    for (double st : step) {
      if ((s = fta_bcs(mesh, 1))) return s;
      ko.dt = st;
      if ((s = hx_compute_fix_therm_admis(mesh, &ko))) return s;
    }
    // the flow's Max_dt (Solver.cpp:1076-1079); only its time step scale remains
    auto cheby = [](double n, double i) { return 1 / (1 - std::cos((n - i - 0.5) * M_PI / n)); };
    const double max_cheby = cheby(o->n_cheby_flow, o->n_cheby_flow - 1);
    ko.use_filter = o->use_filter;
    s = o->use_ldg ? hx_max_dt_navier_stokes(mesh, &ko, o->max_safety / max_cheby, o->max_safety, o->local_time, o->visc,
                                             o->cond, &dummy)
                   : hx_max_dt_euler(mesh, &ko, o->max_safety / max_cheby, o->max_safety, o->local_time, &dummy);
    if (s) return s;
    l = make_launch(mesh);
    HX_TRY(ctx, mesh->ks->fix_av(l, 1), "fix_av");
  }
  --iter;
  *n_iters_out = iter;
  return HX_OK;
}

// Solver::update_art_visc_smoothness (Solver.cpp:457-582) with diffuse_art_visc (Solver.cpp:428-455)
hx_status hx_update_art_visc_smoothness(hx_mesh* mesh, const hx_av_options* o, double* residual)
{
  if (!live(mesh) || !o) return HX_ERR_INVALID;
  hx_ctx* ctx = mesh->ctx;
  const Dev_mesh& dm = mesh->dm;
  for (int c = 0; c < 2; ++c)
    if (dm.n_elem[c] && (!dm.forcing[c] || !dm.advection[c] || !dm.av[c]))
      return fail(ctx, HX_ERR_INVALID, "artificial viscosity needs the av, forcing and advection fields");
  if (!(o->advect_length > 0) || o->advect_iters < 0 || o->diff_iters < 0 || o->n_cheby < 1)
    return fail(ctx, HX_ERR_INVALID, "invalid artificial-viscosity settings");
  if (o->av_rs < 1 || o->av_rs > dm.rs) return fail(ctx, HX_ERR_INVALID, "av_rs must lie in [1, row_size]");
  const double L = o->advect_length;
  const int n_real = 3;  // Storage_params::n_forcing - 1
  const double mult = o->visc_mult * L;
  const double us_max = L * o->unscaled_max * std::sqrt(2 * o->freestream_energy / o->freestream_mass);
  Launch l = make_launch(mesh);
  double* resid_dev = l.scalar + 1;
  auto points = [&](int what) -> hx_status {
    Timer tm(ctx, "set art visc/points", dm.n_elem[0] + dm.n_elem[1]);
    for (int c = 0; c < 2; ++c) HX_TRY(ctx, mesh->ks->av_points(l, c, what, o->av_rs, mult, us_max, resid_dev), "av points");
    return HX_OK;
  };
  hx_status s;
  hx_kernel_options ko{1., 0, 0, 0};
  double dummy;
  // the advection velocity and the Write_face of the scaled state in one pass over the state
  for (int c = 0; c < 2; ++c) {
    Timer tm(ctx, std::string(cls_name(c)) + "/write face/advection velocity", dm.n_elem[c]);
    HX_TRY(ctx, mesh->ks->write_face(l, c, 1), "advection velocity and write_face");
  }
  // enforce CFL condition
  if ((s = hx_compute_prolong(mesh, 0, 0))) return s;
  if ((s = hx_max_dt_advection(mesh, &ko, o->advect_max_safety, 1., 1, L, &dummy))) return s;
  for (int iter = 0; iter < o->advect_iters; ++iter) {
    if ((s = hx_compute_write_face_advection(mesh, L))) return s;
    if ((s = hx_compute_prolong_advection(mesh, L))) return s;
    for (int i = 0; i < 2; ++i) {
      if ((s = hx_apply_advection_bcs(mesh))) return s;
      ko.i_stage = i;
      // the last stage's Locals also form the Legendre projection into forcing[0]
      const bool last = iter + 1 == o->advect_iters && i == 1;
      if ((s = compute_advection_impl(mesh, &ko, L, last ? o->av_rs : 0))) return s;
    }
  }
  // the projection pass, unless the last advection stage formed it
  if (o->advect_iters == 0)
    if ((s = points(1))) return s;
  // root-smear-square
  const double diff_time = o->diff_ratio * L * L / n_real;
  ko = hx_kernel_options{0., 0, 0, 0};
  if ((s = hx_max_dt_smooth_av(mesh, &ko, 1., o->diff_max_safety, 1, &dummy))) return s;
  if ((s = hx_compute_write_face_smooth_av(mesh))) return s;
  if ((s = hx_compute_prolong(mesh, 0, 0))) return s;
  for (int iter = 0; iter < o->diff_iters; ++iter)
    for (int i_cheby = 0; i_cheby < o->n_cheby; ++i_cheby) {
      const double step = 1 / (1 - std::cos((o->n_cheby - i_cheby - 0.5) * M_PI / o->n_cheby));  // math.cpp:104-107
      if ((s = hx_apply_avc_diff_bcs(mesh))) return s;
      ko.dt = step;
      if ((s = hx_compute_smooth_av(mesh, &ko, diff_time, step))) return s;
    }
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, l.stream, resid_dev, 0.);
  HX_TRY(ctx, hipGetLastError(), "fill");
  if ((s = points(2))) return s;
  // the flow state put back (momentum times sqrt(2 rho E)) and its Write_face, in one pass
  for (int c = 0; c < 2; ++c) {
    Timer tm(ctx, std::string(cls_name(c)) + "/write face/restored state", dm.n_elem[c]);
    HX_TRY(ctx, mesh->ks->write_face(l, c, 2), "restore state and write_face");
  }
  if ((s = hx_compute_prolong(mesh, 0, 0))) return s;
  if (residual) {
    double r = 0;
    HX_TRY(ctx, hipMemcpyAsync(&r, resid_dev, sizeof(double), hipMemcpyDeviceToHost, l.stream), "residual readback");
    HX_TRY(ctx, hipStreamSynchronize(l.stream), "residual sync");
    *residual = std::sqrt(r);
  }
  return HX_OK;
}

}  // extern "C"
