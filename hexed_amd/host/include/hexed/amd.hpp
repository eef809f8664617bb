// The device side of the drop-in: a process-wide HIP context and the registry that maps the reference's
// sequences (Kernel_mesh) onto one device mesh.
//
// An Accelerated_mesh walks the sequences of a Kernel_mesh once, assigns every face buffer a slot (face pointers are
// shared between elements and connections, exactly as the reference's connections own the face storage the elements
// point into) and keeps a device copy. Kernels called on registered sequences run on the device copy; host memory is
// a mirror that is refreshed by pull() and written back by push(). Sequences that were never registered take a slow
// path: a temporary device mesh is built, the kernel runs, and the results are scattered back before returning.
#ifndef HEXED_AMD_AMD_HPP_
#define HEXED_AMD_AMD_HPP_

#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../../../include/hexed_amd.h"
#include "Face_permutation.hpp"
#include "Kernel_connection.hpp"
#include "Kernel_mesh.hpp"

namespace hexed::amd
{

hx_ctx* context();
void synchronize();
void check(hx_status s, const char* what);

class Accelerated_mesh
{
  public:
  //! `reference_layout`: element storage follows the reference's Element (time step scale right after the state,
  //! then the bulk and Laplacian artificial viscosity coefficients, the 4 forcing and the row_size advection arrays,
  //! Element.cpp:114-142), so the artificial-viscosity fields are read from it and written back to it
  Accelerated_mesh(Kernel_mesh mesh, bool reference_layout = true);
  ~Accelerated_mesh();
  Accelerated_mesh(const Accelerated_mesh&) = delete;
  Accelerated_mesh& operator=(const Accelerated_mesh&) = delete;
  //! device flow boundary condition for a boundary connection of mesh.def_cons (call before the first push)
  void add_boundary(Kernel_connection& con, hx_bc bc);
  void push();  // host -> device
  void pull();  // device -> host
  hx_mesh* device();
  int n_dim() const { return nd; }
  int row_size() const { return rs; }
  //! registered mesh owning `sequence`, with the element/connection class it holds (-1: the mixed element list)
  static Accelerated_mesh* find(const void* sequence, int* cls);

  // construction from partial lists (the slow path for unregistered sequences)
  Accelerated_mesh(int n_dim, int row_size, std::vector<Kernel_element*> car, std::vector<Kernel_element*> def,
                   std::vector<Kernel_connection*> car_cons, std::vector<Kernel_connection*> def_cons,
                   std::vector<Refined_face*> refs, bool reference_layout, bool register_self);

  private:
  int nd, rs, nq, nfq, nv, face_sz;
  bool ref_layout;
  std::vector<Kernel_element*> elems[2];
  std::vector<Kernel_connection*> cons[2];
  std::vector<Refined_face*> refs;
  std::vector<double*> slot_ptr;   // host address of every face slot (nullptr: the dummy slot)
  std::vector<double*> nslot_ptr;  // host address of every normal slot
  std::unordered_map<double*, int> slot_of, nslot_of;
  std::vector<std::pair<Kernel_connection*, hx_bc>> bcs;
  std::vector<const void*> keys;
  hx_mesh* dev = nullptr;
  // host mirrors in the layout of hx_host_mesh
  std::vector<double> state[2], rcache[2], tss[2], av[2], nom[2], vtss[2], metrics, faces, normals;
  std::vector<double> forcing[2], advection[2];  // reference layout only (Element.cpp:134-142)
  std::vector<int> face_slot[2], face_nslot, con_slot[2], con_dir[2], con_nslot, ref_slot, ref_stretch, bc_con, bc_serial;
  std::vector<hx_bc> bc_defs;
  void build_topology();
  void gather();
  void scatter();
  hx_host_mesh description();
  int slot(double* p);
  int nslot(double* p);
};

//! the device mesh a kernel call on `sequence` runs on: the registered one, or a temporary built for this call
class Mesh_lease
{
  Accelerated_mesh* am = nullptr;
  std::unique_ptr<Accelerated_mesh> temp;
  int c = 0;

  public:
  Mesh_lease(Sequence<Kernel_element&>& s, int n_dim, int row_size, bool deformed);
  Mesh_lease(Sequence<Kernel_connection&>& s, int n_dim, int row_size, bool deformed);
  Mesh_lease(Sequence<Refined_face&>& s, int n_dim, int row_size);
  explicit Mesh_lease(Kernel_mesh& mesh);
  ~Mesh_lease();
  hx_mesh* device() { return am->device(); }
  int cls() const { return c; }
};

//! host reordering of `n_var` variables of face-1 data to match face 0 (Spatial.hpp:73-131), the index map the
//! device's Neighbor uses
std::unique_ptr<Face_permutation_dynamic> face_permutation(int n_dim, int row_size, Connection_direction dir,
                                                           double* data, int n_var);

//! Solver::update (Solver.cpp:834-886) on a registered mesh
void update(Accelerated_mesh& mesh, const hx_update_options& opts, hx_update_status& status);
//! Solver::is_admissible (Solver.cpp:921-959); throws std::runtime_error("state is not finite") as the reference does
bool is_admissible(Accelerated_mesh& mesh);
//! Solver::fix_admissibility (Solver.cpp:961-1095) with the flow settings `opts`; returns its iteration count
int fix_admissibility(Accelerated_mesh& mesh, const hx_update_options& opts, double stability_ratio);
//! Solver::calc_jacobian (Solver.cpp:273-381) on the device from the mesh's vertex data
void calc_jacobian(Accelerated_mesh& mesh, const hx_vertex_data& vertices);

// Device boundary conditions with the arguments and defaults of the reference's Flow_bc constructors
// (Boundary_condition.hpp:61-225). `state` is the conserved freestream state: momentum[n_dim], density, energy.
hx_bc freestream(const std::vector<double>& state);                // Freestream(Mat<>)
hx_bc riemann_invariants(const std::vector<double>& state);        // Riemann_invariants(Mat<>)
hx_bc pressure_outflow(double pressure);                           // Pressure_outflow(double)
hx_bc nonpenetration();                                            // Nonpenetration()
hx_bc copy();                                                      // Copy()
hx_bc outflow();                                                   // Outflow()
//! No_slip(Prescribed_heat_flux(heat_flux), coercion); the reference's default wall is adiabatic with coercion 2
hx_bc no_slip_heat_flux(double heat_flux = 0., double coercion = 2.);
//! No_slip(Prescribed_energy(energy_per_mass), coercion)
hx_bc no_slip_energy(double energy_per_mass, double coercion = 2.);
//! No_slip(Thermal_equilibrium{emissivity, heat_transfer_coef, temperature}, coercion)
hx_bc no_slip_equilibrium(double emissivity, double heat_transfer_coef, double temperature, double coercion = 2.);

}  // namespace hexed::amd
#endif
