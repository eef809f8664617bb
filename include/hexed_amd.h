/*
 * hexed_amd C ABI: the MI355X (gfx950) implementation of hexed's DG spatial discretization and explicit
 * time step, behind plain pointers and sizes.
 *
 * Every kernel entry point replaces one reference kernel class or free function on the hot path; the
 * citation next to each names it (paths relative to the hexed source tree). Adapters that implement the
 * reference's own C++ interfaces (kernel_factory, Spatial<Pde, is_deformed>::Local/Neighbor/..., Solver::update)
 * sit on top of this ABI in hexed_amd/host.
 *
 * Data model. A mesh is described once by `hx_host_mesh` (host arrays, element-major, the reference's
 * per-element layouts) and copied to device memory by `hx_mesh_create`; afterwards all kernels run on the
 * device copy, stream-ordered on the context's stream, and fields move with hx_mesh_upload/hx_mesh_download.
 *
 * Errors. Every call returns an `hx_status`; on failure `hx_last_error` returns the message. Argument
 * errors that the reference raises as std::runtime_error (kernel_factory.hpp:114-116, HEXED_ASSERT in
 * Spatial.hpp:322-323,539-540) return HX_ERR_INVALID with the reference's message text.
 *
 * Threading: one host thread per context; calls on one mesh are not re-entrant (face buffers are shared).
 */
#ifndef HEXED_AMD_H_
#define HEXED_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hx_status {
  HX_OK = 0,
  HX_ERR_INVALID = 1,     /* bad argument (the reference throws std::runtime_error) */
  HX_ERR_HIP = 2,         /* HIP runtime failure */
  HX_ERR_NOT_FINITE = 3,  /* non-finite state detected */
  HX_ERR_UNSUPPORTED = 4, /* combination not built */
  HX_ERR_NO_DEVICE = 5
} hx_status;

/* Limits of kernel_factory (kernel_factory.hpp:111-119; config::max_row_size = 8, CMakeLists.txt:74-77). */
#define HX_MAX_N_DIM 3
#define HX_MIN_ROW_SIZE 2
#define HX_MAX_ROW_SIZE 8

/* Element / connection classes: the reference keeps Cartesian and deformed elements in separate sequences
 * (Kernel_mesh.hpp:14-25). Boundary connections always belong to the deformed class (connection.hpp:321-327). */
#define HX_CARTESIAN 0
#define HX_DEFORMED 1

/* Transport_model (Transport_model.hpp:14-56): coefficient(sqrt_T) =
 * const_val + ref_val*(sqrt_T/sqrt(ref_temp))^3*(ref_temp + temp_offset)/(sqrt_T^2 + temp_offset). */
typedef struct hx_transport {
  double const_val;
  double ref_val;
  double ref_temp;
  double temp_offset;
  int is_viscous;
} hx_transport;

/* The PDE a kernel is instantiated with (the Pde_templ argument of Spatial, pde.hpp). */
typedef enum hx_pde_kind {
  HX_PDE_EULER = 0,          /* pde::Navier_stokes<false>::Pde  (pde.hpp:27-259) */
  HX_PDE_NAVIER_STOKES = 1,  /* pde::Navier_stokes<true>::Pde */
  HX_PDE_ADVECTION = 2,      /* pde::Advection                  (pde.hpp:265-349) */
  HX_PDE_SMOOTH_AV = 3,      /* pde::Smooth_art_visc            (pde.hpp:355-431) */
  HX_PDE_FIX_ADMIS = 4       /* pde::Fix_therm_admis            (pde.hpp:437-493) */
} hx_pde_kind;

/* Constructor arguments of the PDE (the `pde_args...` of every Spatial kernel). */
typedef struct hx_pde {
  int kind;                /* hx_pde_kind */
  hx_transport visc;       /* Navier_stokes: dynamic viscosity */
  hx_transport cond;       /* Navier_stokes: thermal conductivity */
  double advect_length;    /* Advection */
  double diff_time;        /* Smooth_art_visc */
  double cheby_step;       /* Smooth_art_visc */
} hx_pde;

/* Arithmetic form of the flux/derivative stages. HX_FP64 is the reference arithmetic. HX_FP32_MFMA computes
 * the volume fluxes and the interior derivative contractions in FP32 on the matrix cores (FP32 accumulate) and
 * keeps the state, stage accumulation, face-flux lifting terms, metrics and time step in FP64. */
typedef enum hx_precision { HX_FP64 = 0, HX_FP32_MFMA = 1 } hx_precision;

/* Flow boundary conditions applied on device to boundary connections (Boundary_condition.cpp). */
typedef enum hx_bc_kind {
  HX_BC_FREESTREAM = 0,      /* Freestream      (Boundary_condition.cpp:62-76, 299)  */
  HX_BC_COPY = 1,            /* Copy            (Boundary_condition.cpp:450-463)     */
  HX_BC_NONPENETRATION = 2,  /* Nonpenetration  (Boundary_condition.cpp:302-345)     */
  HX_BC_OUTFLOW = 3,         /* Outflow         (Boundary_condition.cpp:465-477)     */
  HX_BC_RIEMANN_INVARIANTS = 4, /* Riemann_invariants (Boundary_condition.cpp:78-182), about `state`; uses
                                   hx_host_mesh::bc_cache */
  HX_BC_PRESSURE_OUTFLOW = 5,   /* Pressure_outflow (Boundary_condition.cpp:184-223), at `pressure`        */
  HX_BC_NO_SLIP = 6             /* No_slip (Boundary_condition.cpp:371-448) with the thermal condition
                                   `thermal`; uses hx_host_mesh::bc_cache */
} hx_bc_kind;

/* Thermal condition of a no-slip wall (Boundary_condition.hpp:140-186). */
typedef enum hx_thermal_kind {
  HX_THERMAL_HEAT_FLUX = 0,   /* Prescribed_heat_flux: wall heat flux `thermal_value` (0: adiabatic, the default) */
  HX_THERMAL_ENERGY = 1,      /* Prescribed_energy: specific energy `thermal_value` at the wall (isothermal)      */
  HX_THERMAL_EQUILIBRIUM = 2  /* Thermal_equilibrium: `emissivity`, `heat_transfer_coef`, `temperature`          */
} hx_thermal_kind;

typedef struct hx_bc {
  int kind;                  /* hx_bc_kind */
  double state[HX_MAX_N_DIM + 2];  /* freestream conserved state: momentum[n_dim], density, energy */
  double pressure;           /* Pressure_outflow: the specified pressure */
  int thermal;               /* No_slip: hx_thermal_kind */
  double thermal_value;      /* No_slip: heat flux or specific energy of the thermal condition */
  double emissivity, heat_transfer_coef, temperature;  /* No_slip with HX_THERMAL_EQUILIBRIUM */
  double coercion;           /* No_slip heat_flux_coercion (Boundary_condition.hpp:198; the reference default is 2) */
} hx_bc;

/*
 * Host description of a mesh. nd = n_dim, rs = row_size, nv = nd + 2, nq = rs^nd, nfq = rs^(nd-1),
 * nvert = 2^nd. Arrays indexed [c] exist per element/connection class (HX_CARTESIAN, HX_DEFORMED); an array
 * may be NULL when its count is 0 or the field is unused (see each field).
 *
 * Layouts follow the reference:
 *  - element fields are [i_elem][i_var][i_qpoint], i_qpoint = (i*rs + j)*rs + k with dimension 0 slowest
 *    (doc/conventions.dox:104-116, Element.cpp:114-194);
 *  - face slots hold `face_sz` doubles; the extrapolated state / flux block is at offset 0 and the LDG
 *    block at offset nv*nfq, each [i_var][i_face_qpoint] (connection.hpp:52-87);
 *    face_sz = max(2*nv, nd + rs)*nfq (connection.hpp:61);
 *  - element face i_face = 2*i_dim + is_positive points at a slot (Element.hpp:44);
 *  - deformed metrics are [i_elem][nd*nd + 1][nq]: reference-level normals [i_dim][j_dim] then the Jacobian
 *    determinant (Deformed_element.cpp:84-90,138-146);
 *  - normal slots hold nd*nfq doubles [i_dim][i_face_qpoint] (connection.hpp:77-86).
 */
typedef struct hx_host_mesh {
  int n_dim;
  int row_size;
  int n_elem[2];
  double* state[2];       /* [n][nv][nq] conserved state: momentum[nd], density, energy       */
  double* rcache[2];      /* [n][max(nv,rs)][nq] stage-0 residual cache (Element.cpp:188; the
                             Advection PDE of the AV generator caches rs variables, Storage_params.cpp:32-35) */
  double* tss[2];         /* [n][nq] local time step scale (Element.cpp:119)                  */
  double* av[2];          /* [n][2][nq] bulk and Laplacian artificial viscosity; NULL => zero */
  double* forcing[2];     /* [n][4][nq] AV forcing (Element.cpp:134); NULL unless AV driver   */
  double* advection[2];   /* [n][rs][nq] AV advection state (Element.cpp:139); NULL unless AV  */
  double* nom_size[2];    /* [n] nominal edge length h (Element.cpp:11)                      */
  double* vertex_tss[2];  /* [n][nvert] vertex time step scale (Element.cpp:163)             */
  int* face_slot[2];      /* [n][2*nd] face slot of each element face                         */
  double* metrics;        /* deformed only: [n][nd*nd + 1][nq]                                */
  int* face_nslot;        /* deformed only: [n][2*nd] normal slot of each face, -1 => identity */
  int n_slot;
  double* faces;          /* [n_slot][face_sz] */
  int n_nslot;
  double* normals;        /* [n_nslot][nd][nfq] */
  int n_con[2];
  int* con_slot[2];       /* [n][2] face slot of side 0 and side 1 */
  int* con_dir[2];        /* [n] packed Connection_direction: i_dim[0] | i_dim[1] << 2 | face_sign[0] << 4 |
                             face_sign[1] << 5 (Kernel_connection.hpp:7-37). Cartesian: {{d, d}, {1, 0}} */
  int* con_nslot;         /* deformed only: [n][2] normal slot of side 0 and side 1 */
  int n_ref;
  int* ref_slot;          /* [n_ref][5]: coarse slot, fine slots 0..3 (-1 when unused) (Refined_face.hpp:9-15) */
  int* ref_stretch;       /* [n_ref][2] */
  int n_bc_con;
  int* bc_con;            /* [n_bc_con] index into the deformed connections */
  int* bc_serial;         /* [n_bc_con] index into bcs */
  int n_bc;
  hx_bc* bcs;
  double* bc_cache;       /* [n_bc_con][nv][nfq] state cache of the boundary faces (Boundary_face::state_cache,
                             written by the state conditions Riemann_invariants and No_slip, read by their flux
                             conditions); may be NULL (the device then keeps its own, zero-initialised) */
} hx_host_mesh;

/* Vertex data of a mesh, which hx_host_mesh does not carry: what Solver::calc_jacobian needs besides the mesh
 * (Solver.cpp:273-381, Deformed_element.cpp:15-136, Hanging_vertex_matcher.hpp:14-33). */
typedef struct hx_vertex_data {
  int n_vertex;                /* number of shared vertices                                                     */
  const int* vertex_id[2];     /* per class [n][2^nd]: the shared vertex of each element vertex, in [0, n_vertex) */
  const double* position;      /* [n_deformed][2^nd][nd] vertex positions of the deformed elements (Vertex::pos)  */
  const double* node_adj;      /* [n_deformed][2 nd][nfq] face node adjustments (Deformed_element::node_adjustments),
                                  [face 2 i_dim + sign][face point]; NULL: all zero (unwarped faces)          */
  int n_hanging;               /* hanging-vertex matchers, one per hanging face                                 */
  const int* hanging;          /* [n_hanging][12]: i_dim, is_positive, stretch[2], then (class, element) of the
                                  2^(nd-1) / (1 + stretch...) fine elements in the matcher's order              */
} hx_vertex_data;

/* Fields that move between host and device (hx_mesh_upload / hx_mesh_download). */
typedef enum hx_field {
  HX_FIELD_STATE = 0,
  HX_FIELD_RCACHE = 1,
  HX_FIELD_TSS = 2,
  HX_FIELD_AV = 3,
  HX_FIELD_FORCING = 4,
  HX_FIELD_ADVECTION = 5,
  HX_FIELD_VERTEX_TSS = 6,
  HX_FIELD_METRICS = 7,   /* class must be HX_DEFORMED */
  HX_FIELD_FACES = 8,     /* class ignored */
  HX_FIELD_NORMALS = 9    /* class ignored */
} hx_field;

/* Kernel_options (kernels.hpp:11-20), minus the stopwatch references. */
typedef struct hx_kernel_options {
  double dt;
  int i_stage;
  int compute_residual;
  int use_filter;
} hx_kernel_options;

/* Solver settings read by Solver::update (Solver.cpp:834-886, defaults at Solver.cpp:151-174). */
typedef struct hx_update_options {
  double max_safety;      /* default 0.7 */
  double max_time_step;   /* default huge */
  int n_cheby_flow;       /* default 1 */
  int flow_iters;         /* default 1 */
  int local_time;         /* local time stepping */
  int use_filter;         /* default 0 */
  int use_ldg;            /* Solver::use_ldg(): viscous transport or artificial viscosity (Solver.cpp:117-120) */
  hx_transport visc;
  hx_transport cond;
  int fix_admis;               /* Solver::set_fix_admissibility, default 0 (Solver.cpp:143, 667-670) */
  double fix_admis_max_safety; /* stability ratio of the fix, default 0.7 (Solver.cpp:153)              */
} hx_update_options;

/* Settings of Solver::update_art_visc_smoothness (Solver.cpp:457-582); the reference's defaults are at
 * Solver.cpp:144,154-163. The mesh needs its forcing and advection arrays (hx_host_mesh::forcing, ::advection). */
typedef struct hx_av_options {
  double advect_length;      /* argument of update_art_visc_smoothness                               */
  int advect_iters;          /* av_advect_iters, default 1                                           */
  double advect_max_safety;  /* av_advect_max_safety, default 0.7                                    */
  double diff_ratio;         /* av_diff_ratio, default 0.3                                           */
  double diff_max_safety;    /* av_diff_max_safety, default 0.7                                      */
  int n_cheby;               /* n_cheby_av, default 1                                                */
  int diff_iters;            /* av_diff_iters, default 1                                             */
  double visc_mult;          /* av_visc_mult, default 1e2                                            */
  double unscaled_max;       /* av_unscaled_max, default 5                                           */
  double freestream_mass;    /* freestream<n_dim>   (density)                                        */
  double freestream_energy;  /* freestream<n_dim+1> (total energy)                                   */
  int av_rs;                 /* Solver::av_rs: degree + 1 of the Legendre projection, default row_size */
} hx_av_options;

/* Status after an update (Iteration_status fields touched by Solver::update, Solver.cpp:872-881). */
typedef struct hx_update_status {
  double time_step;
  double flow_time;
  int iteration;
  int fix_admis_iters;  /* iterations of fix_admissibility (Iteration_status::fix_admis_iters, Solver.cpp:1090) */
} hx_update_status;

typedef struct hx_ctx hx_ctx;
typedef struct hx_mesh hx_mesh;

/* ---- context ---------------------------------------------------------------------------------------- */
hx_status hx_ctx_create(int device, hx_ctx** out);
/* Destroys a context. Meshes of the context that are still alive are detached first: their device memory is freed,
 * their handles stay valid, hx_mesh_destroy on them then only releases the handle, and every other call on them
 * returns HX_ERR_INVALID. Either destruction order is therefore safe; destroying meshes first is the usual one. */
hx_status hx_ctx_destroy(hx_ctx* ctx);
const char* hx_last_error(const hx_ctx* ctx);
hx_status hx_ctx_synchronize(hx_ctx* ctx);
/* The HIP stream (hipStream_t) all work of this context is ordered on. */
void* hx_ctx_stream(hx_ctx* ctx);
/* Number of (n_dim, row_size) combinations built into this library for `pde`; 0 if none. */
int hx_kernels_built(int n_dim, int row_size, int pde_kind);

/* Device timing per kernel category, the device-side counterpart of the reference's Stopwatch_tree categories
 * ("cartesian/local", "deformed/neighbor", "compute time step", "prolong/restrict", "boundary conditions", ...;
 * kernel_factory.hpp:32-45, Solver.cpp:176-201). While enabled, every kernel launch of the context is bracketed by
 * HIP events on the context stream. Category names are "<class>/<kernel>/<pde>", e.g. "deformed/local/navier_stokes".
 * hx_ctx_timing synchronizes the stream, then reports the summed milliseconds, launches and work units (elements or
 * connections processed) of one category; hx_ctx_timing_list writes the category names, '\n'-separated. */
hx_status hx_ctx_set_timing(hx_ctx* ctx, int enable);
hx_status hx_ctx_timing_reset(hx_ctx* ctx);
hx_status hx_ctx_timing(hx_ctx* ctx, const char* category, double* total_ms, long long* launches,
                        long long* work_units);
hx_status hx_ctx_timing_list(hx_ctx* ctx, char* out, size_t capacity);

/* Diagnostic with no counterpart in the reference: a plain device-to-device copy kernel ("hx_stream_copy") over
 * `bytes` bytes (four 16-byte loads in flight per lane before their stores), run `reps` times at each of four grid
 * sizes; returns the best achieved bandwidth counting one read and one write of every byte. Used to measure the
 * achievable HBM bandwidth of the box and to calibrate hardware byte counters. */
hx_status hx_stream_copy(hx_ctx* ctx, size_t bytes, int reps, double* gbs);

/* ---- basis (Basis.hpp:14-63, Gauss_legendre.hpp, Basis.cpp:6-14) ------------------------------------- */
/* Copies one basis table for `row_size` into `out` (capacity `n` doubles). `name` is one of "node", "weight",
 * "diff_mat", "boundary", "orthogonal", "filter", "prolong", "restrict", "eigenvals". Returns the number of
 * doubles the table holds (copying min(n, size)), or -1 for an unknown name / row size. */
int hx_basis_table(int row_size, const char* name, double* out, int n);
/* Basis::max_cfl() = -2*0.9/min_eig_convection (Basis.cpp:6-9) and Basis::step_ratio() = 0.5/0.9 (Basis.cpp:11-14). */
double hx_basis_max_cfl(int row_size);
double hx_basis_step_ratio(int row_size);

/* ---- mesh ------------------------------------------------------------------------------------------- */
/* Copies the topology and every non-NULL field of `desc` to the device. Replaces walking the reference's
 * Kernel_mesh sequences (Kernel_mesh.hpp:14-25) with one registration. Fails with the reference's message
 * "demand for invalid kernel" unless 1 <= n_dim <= 3 and 2 <= row_size <= 8 (kernel_factory.hpp:114-116). */
hx_status hx_mesh_create(hx_ctx* ctx, const hx_host_mesh* desc, hx_mesh** out);
/* Waits for the mesh's pending work and frees it; valid before or after hx_ctx_destroy of its context. */
hx_status hx_mesh_destroy(hx_mesh* mesh);
/* The checks hx_mesh_create applies before touching the device (sizes, every slot / normal / connection index in
 * range, required arrays present); host only. Returns HX_OK or HX_ERR_INVALID with the reason in `msg`. */
hx_status hx_mesh_validate(const hx_host_mesh* desc, char* msg, size_t capacity);
hx_status hx_mesh_upload(hx_mesh* mesh, int field, int cls, const double* host, size_t n_doubles);
hx_status hx_mesh_download(hx_mesh* mesh, int field, int cls, double* host, size_t n_doubles);
/* Arithmetic form of the element kernels (Local) of this mesh; no counterpart in the reference, which is FP64 only.
 * HX_FP64 (default): the reference's arithmetic. HX_FP32_MFMA: volume fluxes, gradients and the derivative
 * contractions in FP32, the contractions on the matrix cores with FP32 accumulation; the state, RK update, residual
 * cache, face fluxes (Neighbor), face jumps, LDG reconciliation, Jacobian determinant scaling and max_dt stay FP64. */
hx_status hx_mesh_set_precision(hx_mesh* mesh, int precision);
/* Device address of a field (for callers that keep their own device views, e.g. torch tensors). hx_update reuses
 * the Max_dt its previous call computed unless the mesh changed in between; a caller that writes through a pointer
 * it obtained before that call must call hx_mesh_device_ptr again (which marks the mesh changed) before the next
 * hx_update. */
hx_status hx_mesh_device_ptr(hx_mesh* mesh, int field, int cls, void** out, size_t* n_doubles);
/* Marks the mesh's device data as changed by the caller (no counterpart in the reference, which has no cached state):
 * call it after writing through a pointer from hx_mesh_device_ptr, so that the next hx_update computes Max_dt afresh
 * instead of reusing the minimum its previous call fused into the last stage. */
hx_status hx_mesh_mark_changed(hx_mesh* mesh);
/* No counterpart in the reference: whether hx_update, hx_update_async and hx_update_no_alloc fold the numerical face
 * fluxes into the Local (each element computes its faces' fluxes from its own and its neighbour's extrapolated state,
 * and Neighbor runs only on connections touching a hanging-face mortar), and likewise the Navier-Stokes jumps into
 * Reconcile_ldg_flux (Neighbor_reconcile only on those connections). mode 1 folds, 0 does not (the reference's
 * sequence), -1 (the default) folds where 2 n_dim <= row_size, where it measured faster. Results agree to rounding
 * (the same expressions); the face state blocks an update leaves are the same, while after a folded Navier-Stokes
 * update the LDG blocks hold the viscous-flux extrapolations instead of the reconciled jumps. Batch and group
 * updates always fold. */
hx_status hx_mesh_set_face_fold(hx_mesh* mesh, int mode);

/* ---- kernels: one per reference kernel class (Spatial.hpp) ------------------------------------------ */
/* Spatial<Pde, def>::Write_face (Spatial.hpp:29-71): extrapolate element state to its faces. */
hx_status hx_write_face(hx_mesh* mesh, int cls, const hx_pde* pde);
/* Spatial<Pde, def>::Neighbor (Spatial.hpp:601-705): numerical flux on every connection of class `cls`. */
hx_status hx_neighbor(hx_mesh* mesh, int cls, const hx_pde* pde, int i_stage);
/* Spatial<Pde, def>::Local (Spatial.hpp:293-510): element-local residual, stage update and Write_face. */
hx_status hx_local(hx_mesh* mesh, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual,
                   int use_filter);
/* Spatial<Pde, def>::Neighbor_reconcile (Spatial.hpp:709-760). */
hx_status hx_neighbor_reconcile(hx_mesh* mesh, int cls, const hx_pde* pde);
/* Spatial<Pde, def>::Reconcile_ldg_flux (Spatial.hpp:513-595). */
hx_status hx_reconcile_ldg_flux(hx_mesh* mesh, int cls, const hx_pde* pde, double dt, int i_stage,
                                int compute_residual, int use_filter);
/* Spatial<Pde, def>::Max_dt (Spatial.hpp:763-829): writes tss and returns the global step in *dt_out
 * (1.0 when is_local). */
hx_status hx_max_dt(hx_mesh* mesh, int cls, const hx_pde* pde, int is_local, int use_filter, double safety_conv,
                    double safety_diff, double* dt_out);
/* Spatial<Pde, false>::Prolong_refined / Restrict_refined (Spatial.hpp:138-286) over all refined faces. */
hx_status hx_prolong_refined(hx_mesh* mesh, const hx_pde* pde, int scale, int offset);
hx_status hx_restrict_refined(hx_mesh* mesh, const hx_pde* pde, int scale, int offset);

/* ---- boundary conditions (Solver.cpp:56-81) --------------------------------------------------------- */
hx_status hx_apply_state_bcs(hx_mesh* mesh);
hx_status hx_apply_flux_bcs(hx_mesh* mesh);

/* ---- free functions (kernels.hpp:22-42) -------------------------------------------------------------- */
/* compute_euler (kernels_convective.cpp:8-18). */
hx_status hx_compute_euler(hx_mesh* mesh, const hx_kernel_options* opts);
/* compute_navier_stokes (kernels_diffusive.cpp:8-29); the flux_bc callback is hx_apply_flux_bcs. */
hx_status hx_compute_navier_stokes(hx_mesh* mesh, const hx_kernel_options* opts, hx_transport visc,
                                   hx_transport cond);
/* max_dt_euler / max_dt_navier_stokes (kernels_max_dt.cpp:8-17): min over both classes. */
hx_status hx_max_dt_euler(hx_mesh* mesh, const hx_kernel_options* opts, double safety_conv, double safety_diff,
                          int local_time, double* dt_out);
hx_status hx_max_dt_navier_stokes(hx_mesh* mesh, const hx_kernel_options* opts, double safety_conv,
                                  double safety_diff, int local_time, hx_transport visc, hx_transport cond,
                                  double* dt_out);
/* compute_write_face / compute_prolong / compute_restrict (kernels_convective.cpp:23-46). */
hx_status hx_compute_write_face(hx_mesh* mesh);
hx_status hx_compute_prolong(hx_mesh* mesh, int scale, int offset);
hx_status hx_compute_restrict(hx_mesh* mesh, int scale, int offset);

/* ---- time step -------------------------------------------------------------------------------------- */
/* Solver::update (Solver.cpp:834-886): flow_iters x n_cheby_flow x [dt <- max_dt; state BCs + stage 0;
 * state BCs + stage 1], each stage followed by fix_admissibility when opts->fix_admis is set (a fix ends the Chebyshev
 * sequence without advancing the time, Solver.cpp:867-870). `status` is advanced in place. */
hx_status hx_update(hx_mesh* mesh, const hx_update_options* opts, hx_update_status* status);
/* Solver::update (Solver.cpp:834-886) n_updates times, enqueued without waiting for the device: each step's time
 * step is formed and kept on the device (the reference passes it by value into every kernel constructor,
 * Solver.cpp:848-849, Spatial.hpp:311), and the status (time_step, flow_time, iteration) accumulates on the device
 * until hx_update_status_sync reads it. `from`, when not NULL, sets the device status first. The first update of a
 * sequence runs Max_dt itself; the steady-state update (starting from the next step's Max_dt that stage 1 fused, or
 * any update with local time stepping) is captured once into a HIP graph and replayed, one graph launch per update,
 * recaptured when the options or the arithmetic form change. Time steps and states are bit-identical to hx_update's.
 * No counterpart of the stage sequence changes; fix_admis (whose verdict is read on the host) and batch meshes are
 * rejected. */
hx_status hx_update_async(hx_mesh* mesh, const hx_update_options* opts, int n_updates, const hx_update_status* from);
/* No counterpart in the reference: hx_update for callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc). The same kernels and results as hx_update, with its scratch -- the
 * time step's reduction scalar and the fused Max_dt minimum of every Euler Local workgroup -- in a file-scope device
 * array of 32 + 262,144 doubles (2 MiB per GPU, held until the process ends) instead of memory the library allocates.
 * Meshes with more than 262,144 Euler Local workgroups (elements / max(1, 256 / row_size^n_dim), both classes) are
 * rejected with HX_ERR_INVALID; fix_admis is rejected with HX_ERR_UNSUPPORTED (the fix allocates its flags). Every
 * call on one GPU shares the array: the fused Max_dt minima a call leaves are used by the mesh's next call only if no
 * other call used the array in between, and calls that overlap on one GPU race. No timing events are recorded. */
hx_status hx_update_no_alloc(hx_mesh* mesh, const hx_update_options* opts, hx_update_status* status);
/* Waits for the mesh's stream and writes the device status of hx_update_async into `status` (time_step, flow_time,
 * iteration; fix_admis_iters is left as it is). */
hx_status hx_update_status_sync(hx_mesh* mesh, hx_update_status* status);
/* Solver::compute_residual (Solver.cpp:899-913): writes the scaled residual into the residual cache. */
hx_status hx_compute_residual(hx_mesh* mesh, const hx_update_options* opts);

/* ---- groups of meshes stepped together ---------------------------------------------------------------------------- */
/* No counterpart in the reference, which runs one Solver per mesh: a group of up to 8 batch meshes (each with its cases
 * registered by hx_mesh_set_cases; a single case is a batch of one) of one n_dim, stepped together. When every mesh is
 * 3D with row size 4, 5 or 6, each kernel of an Euler stage -- the state boundary conditions, Neighbor and Local
 * (Spatial.hpp:293-510, 601-705) -- is launched once for all meshes, both element classes and every row size: a
 * workgroup runs one tile (one per-class workgroup's worth of work) of the (mesh, class) segment it falls in, with the
 * row size as a template argument, so results are bit-identical to stepping each mesh with hx_update_cases. Hanging
 * faces (Prolong/Restrict_refined) and the per-case time steps stay per mesh. Navier-Stokes options, local time stepping
 * and fix_admis step the meshes one at a time. Meshes must outlive the group (hx_mesh_destroy refuses a mesh in a
 * group); hx_ctx_destroy detaches live groups. */
typedef struct hx_group hx_group;
/* A launch shape of the fused Local: order 2 = one tile per workgroup (workgroups_per_cu unused); order 0 = a
 * persistent grid of workgroups_per_cu workgroups per compute unit, tiles strided over the grid; order 1 = the same
 * grid, each workgroup one contiguous chunk of tiles. `ms` is the time the autotuner measured for one launch. */
typedef struct hx_launch_shape {
  int order;
  int workgroups_per_cu;
  double ms;
} hx_launch_shape;
hx_status hx_group_create(hx_ctx* ctx, hx_mesh* const* meshes, int n_meshes, hx_group** out);
hx_status hx_group_destroy(hx_group* group);
/* 1 when the group's Euler stages run as fused launches, 0 when its meshes step one at a time. */
int hx_group_fused(const hx_group* group);
hx_status hx_group_set_shape(hx_group* group, int order, int workgroups_per_cu);
/* One Solver::update of every case of every mesh; `status` holds one entry per case, meshes in order. */
hx_status hx_group_update(hx_group* group, const hx_update_options* opts, hx_update_status* status);
/* Times the fused Local under each candidate launch shape (one tile per workgroup; persistent grids of 2, 4, 8 and 16
 * workgroups per compute unit, strided and chunked), `reps` launches each after a warm-up, the best of them counted, and
 * sets the fastest. The Local runs in residual mode: states are left as they are, the residual cache is overwritten and
 * the face extrapolations go to the second face-state buffer (the primary faces keep theirs; boundary ghost slots get
 * fluxes that the next update's state boundary conditions overwrite), so call it between updates. `table` (up to
 * `capacity` entries, may be NULL) receives every candidate with its time, n_table their number. */
hx_status hx_group_autotune(hx_group* group, int use_filter, int reps, hx_launch_shape* table, int capacity, int* n_table);

/* ---- mesh metrics ------------------------------------------------------------------------------------------------ */
/* Solver::calc_jacobian (Solver.cpp:273-381) without snapping: Element::set_jacobian (Element.cpp:99-112) writes unit
 * face normals of the Cartesian elements and Deformed_element::set_jacobian (Deformed_element.cpp:60-136) the metrics,
 * face normals and vertex time step scale of the deformed ones, from vertex positions (Deformed_element::position,
 * Deformed_element.cpp:15-58); boundary ghost faces copy the inside normal; every deformed connection gets the
 * sign-corrected average of its two element face normals (in normal slots con_nslot, side 1 only when it has a slot
 * of its own); the vertex time step scale is shared by minimum over the elements of each vertex, hanging vertices
 * matched. Cartesian elements start from nominal size / n_dim (Element.cpp:17). The face state blocks hold normals
 * afterwards, as in the reference: write the faces (hx_compute_write_face) before stepping. Hanging faces between
 * deformed elements are not supported (their normals are left as they are). */
hx_status hx_calc_jacobian(hx_mesh* mesh, const hx_vertex_data* vertices);

/* ---- thermodynamic admissibility (Solver.cpp:921-1095; HX_PDE_FIX_ADMIS also runs through the per-kernel entry
 * points above, without compute_residual and the filter) --------------------------------------------------------- */
/* Solver::is_admissible (Solver.cpp:921-959): density and energy positive at every quadrature point of every element,
 * its faces and the fine faces of hanging faces (thermo::admissible, thermo.cpp:6-18). HX_ERR_NOT_FINITE ("state is
 * not finite") when a checked value is not finite. */
hx_status hx_is_admissible(hx_mesh* mesh, int* admissible);
/* Solver::fix_admissibility (Solver.cpp:961-1095) with the flow settings `opts` (its closing Max_dt): while states are
 * inadmissible, and as many iterations again once they are not, two local-time-step diffusion steps of
 * pde::Fix_therm_admis. *n_iters is the reference's return value (0: admissible at once). Afterwards the Laplacian
 * artificial viscosity is zero (the fix coefficient of the last iteration); the mesh's artificial-viscosity field is
 * allocated on the device if it was NULL. */
hx_status hx_fix_admissibility(hx_mesh* mesh, const hx_update_options* opts, double stability_ratio, int* n_iters);
/* compute_fix_therm_admis (kernels_diffusive.cpp:32); the flux_bc callback is the fix's own (ghost LDG = -inside,
 * Solver::apply_fta_flux_bcs, Solver.cpp:103-115). */
hx_status hx_compute_fix_therm_admis(hx_mesh* mesh, const hx_kernel_options* opts);
/* max_dt_fix_therm_admis (kernels_max_dt.cpp:21). */
hx_status hx_max_dt_fix_therm_admis(hx_mesh* mesh, const hx_kernel_options* opts, double msc, double msd,
                                    int local_time, double* dt_out);

/* ---- batch of independent cases ------------------------------------------------------------------------------
 * No counterpart in the reference, which runs one Solver per case: many small cases of one (n_dim, row_size) share
 * one mesh, and each kernel runs once over all of them. */
/* Registers the cases: case_cartesian[i] / case_deformed[i] is the case (0 <= id < n_cases) of element i of each class
 * (NULL for an empty class). Fails unless the cases are independent: no face slot, connection or hanging face may
 * join two cases. */
hx_status hx_mesh_set_cases(hx_mesh* mesh, int n_cases, const int* case_cartesian, const int* case_deformed);
/* Solver::update (Solver.cpp:834-886) of every registered case at once, each case with its own global time step
 * (its own Max_dt minimum, computed and applied on the device); status[n_cases] is advanced in place. Each case's
 * result equals hx_update on a mesh holding that case alone. Afterwards the time step scale field holds each case's
 * last time step (hx_update leaves 1 there). With local time stepping every case takes the same step, as in
 * hx_update. */
hx_status hx_update_cases(hx_mesh* mesh, const hx_update_options* opts, hx_update_status* status);

/* ---- artificial-viscosity generator (the per-kernel entry points above also take HX_PDE_ADVECTION and
 * HX_PDE_SMOOTH_AV; the mesh needs its av, forcing and advection fields) ---------------------------------- */
hx_status hx_apply_advection_bcs(hx_mesh* mesh);                 /* Solver.cpp:508-514, Flow_bc::apply_advection */
hx_status hx_apply_avc_diff_bcs(hx_mesh* mesh);                  /* Solver.cpp:83-91     */
hx_status hx_apply_avc_diff_flux_bcs(hx_mesh* mesh);             /* Solver.cpp:93-101    */
hx_status hx_compute_advection(hx_mesh* mesh, const hx_kernel_options* opts, double advect_length);
                                                                 /* kernels_convective.cpp:8-19 (compute_advection) */
hx_status hx_compute_smooth_av(hx_mesh* mesh, const hx_kernel_options* opts, double diff_time, double cheby_step);
                                                                 /* kernels_diffusive.cpp:8-31 (compute_smooth_av); the
                                                                    flux_bc callback is hx_apply_avc_diff_flux_bcs */
hx_status hx_compute_write_face_advection(hx_mesh* mesh, double advect_length);  /* kernels_convective.cpp:48-51 */
hx_status hx_compute_write_face_smooth_av(hx_mesh* mesh);                         /* kernels_convective.cpp:53-56 */
hx_status hx_compute_prolong_advection(hx_mesh* mesh, double advect_length);     /* kernels_convective.cpp:38-41 */
hx_status hx_max_dt_advection(hx_mesh* mesh, const hx_kernel_options* opts, double msc, double msd, int local_time,
                              double advect_length, double* dt_out);             /* kernels_max_dt.cpp:13-14 */
hx_status hx_max_dt_smooth_av(hx_mesh* mesh, const hx_kernel_options* opts, double msc, double msd, int local_time,
                              double* dt_out);                                   /* kernels_max_dt.cpp:15 */
/* Solver::update_art_visc_smoothness (Solver.cpp:457-582): writes the bulk coefficient av[.][0] and, if `residual`
 * is not NULL, the art_visc_residual it assigns (Solver.cpp:575). */
hx_status hx_update_art_visc_smoothness(hx_mesh* mesh, const hx_av_options* opts, double* residual);
/* Stab_art_visc (stabilizing_art_visc.cpp:8-61; the stabilizing_art_visc free function over a Kernel_mesh): per
 * element the specific volume's share of L2 energy in the top Legendre mode along each dimension, on a sine ramp of
 * half width 0.5 around log10 = -4.25 log10(row_size - 1), times (row_size - 1) char_speed nominal_size, written to
 * uncert[n_cartesian + n_deformed] (Kernel_element::uncert(); Cartesian elements first). */
hx_status hx_stab_art_visc(hx_mesh* mesh, double char_speed, double* uncert);
/* Solver::set_art_visc_admis (Solver.cpp:636-659) with char_speed = freestream speed + freestream sound speed:
 * Stab_art_visc's element values on every vertex, shared by maximum over the elements of each vertex (hanging
 * vertices matched, as in hx_calc_jacobian; `vertices` needs the ids and matchers, not the positions), interpolated
 * n-linearly to the quadrature points into the Laplacian coefficient av[.][1]. `uncert`, when not NULL, receives the
 * element values. The mesh keeps the vertex ids and matchers of its last call here or to hx_calc_jacobian on the device,
 * with a host copy: a call whose topology equals it (compared by content, not by pointer) neither validates nor uploads
 * them again, and the work arrays stay allocated, so repeated calls (Case.cpp:432 calls it before every update) cost
 * the kernels and one readback of `uncert`. */
hx_status hx_set_art_visc_admis(hx_mesh* mesh, const hx_vertex_data* vertices, double char_speed, double* uncert);

#ifdef __cplusplus
}
#endif
#endif
