// gks_host.h — the reference solver's 1D, 2D and 3D call surface (types,
// globals, function-pointer plugin points and sweep functions) over the
// gfx950 C ABIs (include/gks1d.h, gks2d.h, gks3d.h).  A problem driver
// written against the reference (src/1Dproblem.cpp, 2Dproblem.cpp,
// 3Dproblem.cpp style: set globals, assign the scheme pointers, call the
// sweeps in order) compiles against these headers unchanged and runs its
// per-step hot path on the GPU.
//
// Names, argument meanings and error behaviour follow the reference headers
// (include/fluid_mesh.h, basic_function.h, boundary_condition.h,
// reconstruction.h, flux_function.h, time_advance.h, output.h).  What
// differs, by design:
//  * face records (Interface2d gauss states, Flux2d_gauss fluxes) live on the
//    device; the host arrays Setinterface_array / Setflux_gauss_array return
//    carry geometry only;
//  * the cell array is uploaded at the first sweep that touches it and copied
//    back by output2d() or Download_fluid(); host writes to `fluids` between
//    sweeps need Upload_fluid();
//  * a blow-up found by Convar_to_Primvar is reported (message + exit(0), as
//    the reference does) at the next Get_CFL;
//  * the record types carry every field of the reference's; the fields the
//    device sweeps keep on the device (Interface2d::line / gauss,
//    Flux2d_gauss::gauss, Fluid2d's stage and HWENO work arrays) exist on
//    the host with storage but are not filled by the sweeps;
//  * which kernel build a context uses is the shim global gks_fast_kernels
//    (false: the IEEE parity build; true: the contracted fast build).
#pragma once
#include <cmath>
#include <ctime>
#include <cstring>
#include <iostream>
#include <string>

using namespace std;

// ---------------------------------------------------------------- fluid_mesh
extern int gausspoint;
extern double* gauss_loc;
extern double* gauss_weight;

class Block2d {
 public:
  int index = 0;
  bool uniform = true;
  int ghost = 5;
  int nodex = 0, nodey = 0;
  int nx = 0, ny = 0;
  int xcell_begin = 0, xcell_end = 0, ycell_begin = 0, ycell_end = 0;
  int xinterface_begin_n = 0, xinterface_begin_t = 0, xinterface_end_n = 0, xinterface_end_t = 0;
  int yinterface_begin_n = 0, yinterface_begin_t = 0, yinterface_end_n = 0, yinterface_end_t = 0;
  double dx = 0, dy = 0, overdx = 0, overdy = 0;
  double left = 0, right = 0, down = 0, up = 0;
  int stages = 0;
  double timecoefficient[5][5][2] = {};
  double t = 0;
  double CFL = 0;
  double dt = 0;
  int step = 0;
  double T = 0;
  int outputperstep = 0;
};

class Point2d {
 public:
  double convar[4];
  double primvar[4];
  double der1x[4];
  double der1y[4];
  double x, y;
  double normal[2];
  bool is_reduce_order;
};

// fluid_mesh.h:141-171, every field; the sweeps read and write convar,
// alpha_x / alpha_y and (for Get_CFL at step 0) primvar
class Fluid2d {
 public:
  double primvar[4];
  double exact;
  double convar[4];
  double convar_old[4];
  double convar_stage1[4];
  double convar_stage2[2];
  double res[4];
  double xindex;
  double yindex;
  double coordx;
  double coordy;
  double dx;
  double dy;
  double area;
  bool is_hweno;
  double node[8];
  bool boundary;
  double alpha_x = 0.0;
  double alpha_y = 0.0;
  string direction;
  bool is_left = false;
  double Sw1[2][4];
  double Lw1[2][4];
  double derSw1[2][4];
  double derLw1[2][4];
};

class Flux2d {
 public:
  double x[4];
  double f[4];
  double derf[4];
  double length;
  double normal[2];
};

class Recon2d {
 public:
  Point2d left, center, right;
  double x, y;
  double normal[2];
};

// fluid_mesh.h:194-206: geometry of one face; its reconstructed states
// (line, gauss) are computed and kept on the device: the host records
// (gausspoint of them per face, allocated by Setinterface_array) stay zero
class Interface2d {
 public:
  Recon2d line;
  Recon2d* gauss = nullptr;
  double x = 0, y = 0;
  double normal[2] = {0, 0};
  double length = 0;
};

// fluid_mesh.h:208-216: per-face flux records (gausspoint per face,
// allocated by Setflux_gauss_array); the device keeps the fluxes, the host
// records stay zero
class Flux2d_gauss {
 public:
  Flux2d* gauss = nullptr;
};

void SetGuassPoint();

// ------------------------------------------------------------ basic_function
extern int K;
extern double Gamma;
extern double Mu;
extern double Nu;
extern double T_inf;
extern double c1_euler;
extern double c2_euler;
extern bool is_Prandtl_fix;
extern double Pr;
extern double R_gas;
enum TAU_TYPE { Euler, NS };
extern TAU_TYPE tau_type;

Fluid2d* Setfluid(Block2d& block);
Interface2d* Setinterface_array(Block2d block);
Flux2d_gauss** Setflux_gauss_array(Block2d block);
void SetUniformMesh(Block2d& block, Fluid2d* fluids, Interface2d* xinterfaces, Interface2d* yinterfaces,
                    Flux2d_gauss** xfluxes, Flux2d_gauss** yfluxes);
void CopyFluid_new_to_old(Fluid2d* fluids, Block2d block);
void Convar_to_Primvar(Fluid2d* fluids, Block2d block);
void Copy_Array(double* target, double* origin, int dim);
void Convar_to_primvar_2D(double* primvar, double convar[4]);
void Primvar_to_convar_2D(double* convar, double primvar[4]);

// -------------------------------------------------------- boundary_condition
typedef void (*BoundaryCondition2d)(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void free_boundary_left(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void free_boundary_right(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void free_boundary_down(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void free_boundary_up(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void periodic_boundary_left(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void periodic_boundary_right(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void periodic_boundary_down(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void periodic_boundary_up(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void reflection_boundary_left(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void reflection_boundary_right(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void reflection_boundary_down(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);
void reflection_boundary_up(Fluid2d* fluids, Block2d block, Fluid2d bcvalue);

// ------------------------------------------------------------ reconstruction
enum Reconstruction_variable { conservative, characteristic };
extern Reconstruction_variable reconstruction_variable;
enum WENOtype { linear, wenojs, wenoz };
extern WENOtype wenotype;
extern bool is_reduce_order_warning;
extern bool is_using_df_factor;

typedef void (*Reconstruction_within_Cell_2D_normal)(Interface2d& left, Interface2d& right, Interface2d& down,
                                                     Interface2d& up, Fluid2d* fluids, Block2d block);
extern Reconstruction_within_Cell_2D_normal cellreconstruction_2D_normal;
typedef void (*Reconstruction_within_Cell_2D_tangent)(Interface2d* left, Interface2d* right, Interface2d* down,
                                                      Interface2d* up, Fluid2d* fluids, Block2d block);
extern Reconstruction_within_Cell_2D_tangent cellreconstruction_2D_tangent;
typedef void (*Reconstruction_forG0_2D_normal)(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids,
                                               Block2d block);
extern Reconstruction_forG0_2D_normal g0reconstruction_2D_normal;
typedef void (*Reconstruction_forG0_2D_tangent)(Interface2d* xinterfaces, Interface2d* yinterfaces,
                                                Fluid2d* fluids, Block2d block);
extern Reconstruction_forG0_2D_tangent g0reconstruction_2D_tangent;

// plugin leaves: assignable to the pointers above; the sweeps dispatch on
// their identity (a leaf called directly is an error: its work is fused
// into the device sweep)
void WENO5_AO_normal(Interface2d& left, Interface2d& right, Interface2d& down, Interface2d& up, Fluid2d* fluids,
                     Block2d block);
void WENO5_AO_with_df_normal(Interface2d& left, Interface2d& right, Interface2d& down, Interface2d& up,
                             Fluid2d* fluids, Block2d block);
void WENO5_AO_tangent(Interface2d* left, Interface2d* right, Interface2d* down, Interface2d* up, Fluid2d* fluids,
                      Block2d block);
void WENO5_AO_with_df_tangent(Interface2d* left, Interface2d* right, Interface2d* down, Interface2d* up,
                              Fluid2d* fluids, Block2d block);
// WENO7 / WENO9 with the DF factor (reconstruction.h:62-80); need
// gausspoint 3 / 4
void WENO7_AO_with_df_normal(Interface2d& left, Interface2d& right, Interface2d& down, Interface2d& up,
                             Fluid2d* fluids, Block2d block);
void WENO7_AO_with_df_tangent(Interface2d* left, Interface2d* right, Interface2d* down, Interface2d* up,
                              Fluid2d* fluids, Block2d block);
void WENO9_AO_with_df_normal(Interface2d& left, Interface2d& right, Interface2d& down, Interface2d& up,
                             Fluid2d* fluids, Block2d block);
void WENO9_AO_with_df_tangent(Interface2d* left, Interface2d* right, Interface2d* down, Interface2d* up,
                              Fluid2d* fluids, Block2d block);
void Center_do_nothing_normal(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids, Block2d block);
void Center_all_collision_multi(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids,
                                Block2d block);

void Reconstruction_within_cell(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids,
                                Block2d block);
void Reconstruction_forg0(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids, Block2d block);
void Update_alpha(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids, Block2d block);

// ------------------------------------------------------------- flux_function
enum GKS2d_type { nothing_2d, kfvs1st_2d, kfvs2nd_2d, gks1st_2d, gks2nd_2d, pp_gks_2d };
extern GKS2d_type gks2dsolver;
typedef void (*Flux_function_2d)(Flux2d& flux, Recon2d& re, double dt);
extern Flux_function_2d flux_function_2d;
// one point on the GPU: centre state + GKS2D (flux_function.cpp:345-698)
void GKS2D(Flux2d& flux, Recon2d& interface, double dt);
void Calculate_flux(Flux2d_gauss** xfluxes, Flux2d_gauss** yfluxes, Interface2d* xinterfaces,
                    Interface2d* yinterfaces, Block2d block, int stage);

// -------------------------------------------------------------- time_advance
typedef void (*TimeMarchingCoefficient_2d)(Block2d& block);
extern TimeMarchingCoefficient_2d timecoe_list_2d;
void S2O4_2D(Block2d& block);
void Initial_stages(Block2d& block);
double Get_CFL(Block2d& block, Fluid2d* fluids, double tstop);
void Update(Fluid2d* fluids, Flux2d_gauss** xfluxes, Flux2d_gauss** yfluxes, Block2d block, int stage);

// -------------------------------------------------------------------- output
struct Runtime {
  clock_t start_initial = 0, finish_initial = 0, start_compute = 0, finish_compute = 0;
};
// writes convar[0] of the interior, j outer / i inner, to ../data/R-<step>.txt
// (output.cpp:64-84), after copying the device state back
void output2d(Fluid2d* fluids, Block2d block);

// ----------------------------------------------------------- problem set-up
void IC_for_riemann_2d(Fluid2d* fluid, double* zone1, double* zone2, double* zone3, double* zone4, Block2d block,
                       double* discon);

// --------------------------------------------- device residency (shim only)
// kernel build of the device contexts created from now on (1D, 2D, 3D):
// false (default) the IEEE build, whose results differ from the reference
// only through the device erfc / exp / sqrt; true the fast build (FMA
// contraction, reciprocal multiplication)
extern bool gks_fast_kernels;
// copies the device state (convar, alpha) back into `fluids` and refreshes
// primvar on the host
void Download_fluid(Fluid2d* fluids, Block2d block);
// pushes host-side edits of `fluids` (convar, alpha, primvar for the next
// Get_CFL) to the device
void Upload_fluid(Fluid2d* fluids, Block2d block);
// releases the device state bound to `fluids`
void Release_fluid(Fluid2d* fluids);

// ======================================================================= 3D
// (fluid_mesh.h:225-351, basic_function.h:585-710, boundary_condition.h:35-41,
// reconstruction.h:96-136, flux_function.h:30-41, time_advance.h:47-53,
// output.h:33).  The 3D scheme is the one of include/gks3d.h: the reference's
// algorithm with lambda from the state (its GKS3D takes the pressure) and the
// Gauss-point records on the device (it never allocates them).
extern double** gauss_loc_3d;
extern bool is_multi_dimensional;

class Block3d {
 public:
  bool uniform = true;
  int ghost = 3;
  int nodex = 0, nodey = 0, nodez = 0;
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0;
  double overdx = 0, overdy = 0, overdz = 0;
  double xarea = 0, yarea = 0, zarea = 0;
  double left = 0, right = 0;
  double down = 0, up = 0;
  double front = 0, back = 0;
  int stages = 0;
  double timecoefficient[5][5][3] = {};
  double timecoefficient_hweno[5][5][3] = {};
  double t = 0;
  double CFL = 0;
  double dt = 0;
  double tstop = 0;
  int step = 0;
  double xcell_begin = 0, xcell_end = 0, ycell_begin = 0, ycell_end = 0, zcell_begin = 0, zcell_end = 0;
  int index = 0;
};

class Fluid3d {
 public:
  double primvar[6];
  double convar[6];
  double convar_old[6];
  int index[3];
  double loc[3];
  double node[24];
  double vol;
  double R_c;
};

class Point3d {
 public:
  double convar[6];
  double der1x[6];
  double der1y[6];
  double der1z[6];
};

class Recon3d {
 public:
  Point3d left;
  Point3d center;
  Point3d right;
  bool order_reduce;
};

// geometry of one face; its line and Gauss-point states live on the device
class Interface3d {
 public:
  double loc[12] = {};
  double weight[4] = {};
  double normal[12] = {};
  int index[3] = {0, 0, 0};
  double area = 0;
};

class Flux3d {
 public:
  double x[6];
  double f[6];
  double derf[6];
};

// per-face flux records live on the device
class Flux3d_gauss {
 public:
  Flux3d* gauss = nullptr;
};

void SetGuassPoint_3D();

Fluid3d* Setfluid_array(Block3d& block);
Interface3d* Setinterface_array(Block3d& block);
Flux3d_gauss** Setflux_gauss_array(Block3d& block);
void SetUniformMesh(Block3d block, Fluid3d* fluids, Interface3d* xinterfaces, Interface3d* yinterfaces,
                    Interface3d* zinterfaces);
void CopyFluid_new_to_old(Fluid3d* fluids, Block3d block);
void Convar_to_primvar(Fluid3d* fluids, Block3d& block);
void Convar_to_primvar_3D(double* primvar, double convar[5]);
void Primvar_to_convar_3D(double* convar, double* primvar);

typedef void (*BoundaryCondition3d)(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_xleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_xright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_yleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_yright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_zleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void free_boundary_zright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
// periodic 3D boundaries: not in the reference, same signature
void periodic_boundary_xleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void periodic_boundary_xright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void periodic_boundary_yleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void periodic_boundary_yright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void periodic_boundary_zleft(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);
void periodic_boundary_zright(Fluid3d* fluids, Block3d block, Fluid3d bcvalue);

enum G0_construct_type { collisionn, collisionnless, all_collisionn };
extern G0_construct_type g0type;
typedef void (*Reconstruction_within_Cell_3D_of_face_value)(Interface3d& left, Interface3d& right, Interface3d& down,
                                                            Interface3d& up, Interface3d& back, Interface3d& front,
                                                            Fluid3d* fluids, Block3d block);
extern Reconstruction_within_Cell_3D_of_face_value cellreconstruction_3D_of_face_value;
typedef void (*Reconstruction_within_Cell_3D_of_line_value)(Interface3d* xinterface, Interface3d* yinterface,
                                                            Interface3d* zinterface, Block3d block);
extern Reconstruction_within_Cell_3D_of_line_value cellreconstruction_3D_of_line_value;
typedef void (*Reconstruction_within_Cell_3D_of_point_value)(Interface3d* xinterface, Interface3d* yinterface,
                                                             Interface3d* zinterface, Block3d block);
extern Reconstruction_within_Cell_3D_of_point_value cellreconstruction_3D_of_point_value;
typedef void (*Reconstruction_forG0_3D_of_face_value)(Interface3d* xinterfaces, Interface3d* yinterfaces,
                                                      Interface3d* zinterfaces, Fluid3d* fluids, Block3d block);
extern Reconstruction_forG0_3D_of_face_value g0reconstruction_3D_of_face_value;
typedef void (*Reconstruction_forG0_3D_of_line_value)(Interface3d* xinterfaces, Interface3d* yinterfaces,
                                                      Interface3d* zinterfaces, Block3d block);
extern Reconstruction_forG0_3D_of_line_value g0reconstruction_3D_of_line_value;
typedef void (*Reconstruction_forG0_3D_of_point_value)(Interface3d* xinterfaces, Interface3d* yinterfaces,
                                                       Interface3d* zinterfaces, Block3d block);
extern Reconstruction_forG0_3D_of_point_value g0reconstruction_3D_of_point_value;

// plugin leaves (their work is fused into the device sweeps)
void WENO5_AO_splitting_3d(Interface3d& left, Interface3d& right, Interface3d& down, Interface3d& up,
                           Interface3d& back, Interface3d& front, Fluid3d* fluids, Block3d block);
void WENO5_AO_of_line_value(Interface3d* right, Interface3d* up, Interface3d* front, Block3d block);
void WENO5_AO_of_point_value(Interface3d* right, Interface3d* up, Interface3d* front, Block3d block);
void Do_nothing_splitting_3d(Interface3d* xinterfaces, Interface3d* yinterfaces, Interface3d* zinterfaces,
                             Fluid3d* fluids, Block3d block);
void Do_nothing_reconstruction_of_line_value(Interface3d* xinterfaces, Interface3d* yinterfaces,
                                             Interface3d* zinterfaces, Block3d block);
void Do_nothing_reconstruction_of_point_value(Interface3d* xinterfaces, Interface3d* yinterfaces,
                                              Interface3d* zinterfaces, Block3d block);
void Reconstruction_within_cell(Interface3d* xinterfaces, Interface3d* yinterfaces, Interface3d* zinterfaces,
                                Fluid3d* fluids, Block3d block);
void Reconstruction_forg0(Interface3d* xinterfaces, Interface3d* yinterfaces, Interface3d* zinterfaces,
                          Fluid3d* fluids, Block3d block);

enum GKS3d_type { nothing_3d, kfvs1st_3d, kfvs2nd_3d, gks1st_3d, gks2nd_3d };
extern GKS3d_type gks3dsolver;
typedef void (*Flux_function_3d)(Flux3d& flux, Recon3d& re, double area, double dt);
extern Flux_function_3d flux_function_3d;
// one Gauss point on the GPU (gks2nd_3d with the S2O4 pair)
void GKS3D(Flux3d& flux, Recon3d& interface, double area, double dt);
void Calculate_flux(Flux3d_gauss** xfluxes, Flux3d_gauss** yfluxes, Flux3d_gauss** zfluxes, Interface3d* xinterfaces,
                    Interface3d* yinterfaces, Interface3d* zinterfaces, Block3d block, int stage);

typedef void (*TimeMarchingCoefficient_3d)(Block3d& block);
extern TimeMarchingCoefficient_3d timecoe_list_3d;
void S2O4_3D(Block3d& block);
void Initial_stages(Block3d& block);
double Get_CFL(Block3d& block, Fluid3d* fluids, double tstop);
void Update_gauss(Fluid3d* fluids, Flux3d_gauss** xfluxes, Flux3d_gauss** yfluxes, Flux3d_gauss** zfluxes,
                  Block3d block, int stage);

// writes convar[0] of the interior, k outer / j / i inner, to ../data/R3d.txt
// (output.cpp:86-104), after copying the device state back
void output3d(Fluid3d* fluids, Block3d block);

// ICfor_cubic_sod (3Dproblem.cpp:174-223)
void ICfor_cubic_sod(Fluid3d* fluids, double* zone1, double* zone2, Block3d block);

// device residency (shim only), as for Fluid2d
void Download_fluid(Fluid3d* fluids, Block3d block);
void Upload_fluid(Fluid3d* fluids, Block3d block);
void Release_fluid(Fluid3d* fluids);

// ======================================================================= 1D
// (fluid_mesh.h:8-75, basic_function.h:587-689, boundary_condition.h:5-11,
// reconstruction.h:14-44, flux_function.h:6-13, time_advance.h:5-20,
// output.h:27, 1Dproblem.h:10)
class Block1d {
 public:
  int ghost = 5;
  int nodex = 0;
  int nx = 0;
  int nodex_begin = 0;
  int nodex_end = 0;
  double dx = 0;
  double left = 0;
  double right = 0;
  int stages = 0;
  double timecoefficient[5][5][2] = {};
  double t = 0;
  double CFL = 0;
  double dt = 0;
  int step = 0;
};

class Fluid1d {
 public:
  double primvar[3];
  double convar[3];
  double convar_old[3];
  double convar_stage1[3];
  double convar_stage2[3];
  double cx;
  double dx;
  double alpha = 0.0;
};

class Point1d {
 public:
  double convar[3];
  double convar_old[3];
  double der1[3];
  double x;
};

class Flux1d {
 public:
  double F[3];
  double f[3];
  double derf[4];
};

// the face states of the device sweeps stay on the device; the host records
// keep geometry, and GKS() fills left / center / right of one record
class Interface1d {
 public:
  Point1d left;
  Point1d center;
  Point1d right;
  Flux1d* flux = nullptr;
  double x = 0;
};

Flux1d** Setflux_array(Block1d block);
void SetUniformMesh(Block1d block, Fluid1d* fluids, Interface1d* interfaces, Flux1d** fluxes);
void CopyFluid_new_to_old(Fluid1d* fluids, Block1d block);
void Convar_to_primvar_1D(double* primvar, double convar[3]);
void Primvar_to_convar_1D(double* convar, double primvar[3]);

typedef void (*BoundaryCondition)(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void free_boundary_left(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void free_boundary_right(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void reflection_boundary_left(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void reflection_boundary_right(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void periodic_boundary_left(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);
void periodic_boundary_right(Fluid1d* fluids, Block1d block, Fluid1d bcvalue);

typedef void (*Reconstruction_within_Cell)(Point1d& left, Point1d& right, Fluid1d* fluids, Block1d block);
extern Reconstruction_within_Cell cellreconstruction;
typedef void (*Reconstruction_forG0)(Interface1d& interfaces, Fluid1d* fluids);
extern Reconstruction_forG0 g0reconstruction;
// plugin leaves (their work is fused into the device sweeps)
void WENO5_AO(Point1d& left, Point1d& right, Fluid1d* fluids, Block1d block);
void WENO5_AO_with_DF(Point1d& left, Point1d& right, Fluid1d* fluids, Block1d block);
void Center_collision(Interface1d& interfaces, Fluid1d* fluids);
void Reconstruction_within_cell(Interface1d* interfaces, Fluid1d* fluids, Block1d block);
void Reconstruction_forg0(Interface1d* interfaces, Fluid1d* fluids, Block1d block);
void Update_alpha(Interface1d* interfaces, Fluid1d* fluids, Block1d block);

enum GKS1d_type { nothing, kfvs1st, kfvs2nd, gks1st, gks2nd, pp_gks };
extern GKS1d_type gks1dsolver;
typedef void (*Flux_function)(Flux1d& flux, Interface1d& interface, double dt);
extern Flux_function flux_function;
// one interface on the GPU: Center_collision + GKS (flux_function.cpp:16-276)
void GKS(Flux1d& flux, Interface1d& interface, double dt);
void Calculate_flux(Flux1d** fluxes, Interface1d* interfaces, Block1d& block, int stage);

typedef void (*TimeMarchingCoefficient)(Block1d& block);
extern TimeMarchingCoefficient timecoe_list;
void S2O4(Block1d& block);
void Initial_stages(Block1d& block);
double Get_CFL(Block1d& block, Fluid1d* fluids, double tstop);
void Update(Fluid1d* fluids, Flux1d** fluxes, Block1d block, int stage);

// writes convar[0] of the interior to ../data/R.txt (output.cpp:3-15), after
// copying the device state back
void output1d(Fluid1d* fluids, Block1d block);
// ICforSodTube (1Dproblem.cpp:150-170)
void ICforSodTube(Fluid1d* fluids, Block1d block);

// device residency (shim only)
void Download_fluid(Fluid1d* fluids, Block1d block);
void Upload_fluid(Fluid1d* fluids, Block1d block);
void Release_fluid(Fluid1d* fluids);
