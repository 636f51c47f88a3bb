/*
 * ocfd.h -- C ABI of the MI355X-native OpenCFD-AMR per-box advance.
 *
 * Plain C: POD structs, device pointers as double*, int status codes
 * (0 = ok), no exceptions and no torch types.  A "level" is the drop-in for one
 * AMReX level's BoxArray + Geometry + BCs: it owns the device box work-lists
 * and lays out MultiFab storage as one contiguous device buffer holding every
 * box's FAB (component-major SoA, i fastest, boxes in list order), which is
 * the amrex::Array4 convention of the reference (Source/Advance.cpp:224-228).
 *
 * Every entry point cites the reference interface it replaces.
 */
#ifndef OCFD_H
#define OCFD_H

#ifdef __cplusplus
extern "C" {
#endif

#define OCFD_NCONS 5   /* Source/IndexDefines.H:4-9   URHO,UMX,UMY,UMZ,UEDEN */
#define OCFD_NSPEC 9   /* Source/CHEM/LiDryer.H       H2,O2,H2O,H,O,OH,HO2,H2O2,N2 */
#define OCFD_NGROW 4   /* Source/EBR.H:207            NUM_GROW              */

/* status codes */
#define OCFD_OK           0
#define OCFD_EINVAL       1
#define OCFD_EHIP         2
#define OCFD_ENOMEM       3
#define OCFD_EUNSUPPORTED 4
#define OCFD_ENONFINITE   5   /* a state holds a non-finite value (NaN or inf) */

/* flags for the flux entry points */
#define OCFD_VISC       1   /* ebr.do_visc    (Source/EBR.cpp:25)  */
#define OCFD_GRAVITY    2   /* ebr.do_gravity (Source/EBR.cpp:26)  */
#define OCFD_FLUX_FP32  4   /* convective/viscous fluxes in FP32, state and accumulation FP64 */
#define OCFD_PLANAR     8   /* z-uniform data: skip the z sweep (its divergence is exactly 0) */
#define OCFD_EB_REDIST 16   /* ebr.do_redistribute (Source/EBR.cpp:27): small cut cells scaled by vfrac */

/* struct Parm, Source/Constants.H:7-64 (non-CHEM members, same order) */
typedef struct ocfd_parm {
    double eos_gamma, eos_m, Patm, cv, cp, Ru, Rg, kOverMu, Pr, C_s, T_s,
           smallr, smallp, eps, twoThirds, oneSix;
} ocfd_parm;

/* Level geometry + physical BCs: amrex::Geometry (domain, prob_lo, dx,
 * periodicity) + ebr.lo_bc/hi_bc codes (0 Interior, 1 Inflow, 2 Outflow,
 * 3 Symmetry, 4 SlipWall, 5 NoSlipWall; Exec/<case>/inputs) + the case's
 * ext_dir fill (probtype 1 = Exec/1_DMR/prob_bcfill.cpp; probtype 2 = the inflow
 * state left_state at x-lo, Exec/3_FSTEP/prob_bcfill.cpp). */
typedef struct ocfd_geom {
    int dom_lo[3], dom_hi[3];
    double prob_lo[3];
    double dx[3];
    int periodic[3];
    int lo_bc[3], hi_bc[3];
    int probtype;
    double left_state[5];
    double right_state[5];
} ocfd_geom;

typedef struct ocfd_level_s* ocfd_level;
typedef struct ocfd_graph_s* ocfd_graph;   /* captured device-driven steps */

/* Embedded boundary: a body extruded along z, given by an implicit function
 * (> 0 inside the body) -- the eb2.geom_type choices the cases use
 * (Exec/3_FSTEP/inputs box, Exec/4_WEDGE/inputs plane) plus a cylinder.
 * fluid_inside flips the sign (fluid inside the cylinder or box).
 * Cells whose volume fraction is below small_volfrac (eb2.small_volfrac;
 * <= 0 means 1e-14) are covered. */
#define OCFD_EB_NONE     0
#define OCFD_EB_CYLINDER 1   /* centre (center[0], center[1]), radius */
#define OCFD_EB_BOX      2   /* [lo, hi] in x and y */
#define OCFD_EB_PLANE    3   /* body where (x - point) . normal > 0 (x and y components) */
typedef struct ocfd_eb_spec {
    int kind;
    int fluid_inside;
    double center[3];
    double radius;
    double lo[3], hi[3];
    double point[3], normal[3];
    double small_volfrac;
} ocfd_eb_spec;

/* Device-resident time control of one level: the state of the coarse time
 * loop (Source/main.cpp:61-70 with EBR::computeInitialDt / computeNewDt,
 * Source/EBR.cpp:189-306), kept in device memory so a whole time step --
 * ghost fills, RK stages, CFL reduction and the dt update -- runs without a
 * host round trip and can be captured into a hipGraph.  Allocate
 * sizeof(ocfd_step_ctl) (256) bytes of 128-byte aligned device memory per
 * level. */
typedef struct ocfd_step_ctl {
    double time;            /* current time */
    double dt;              /* dt of the next step */
    double cfl;
    double stop_time;       /* < 0: none */
    int steps;              /* steps taken */
    int max_steps;          /* < 0: none */
    int done;               /* set when stop_time or max_steps is reached, or on a non-finite state */
    int status;             /* OCFD_OK, or OCFD_ENONFINITE once a step produced a non-finite cell */
    char pad1[80];
    /* CFL reduction slot (bit pattern of a positive double) and the
     * non-finite flag of the reduction, on a cache line of their own: they
     * are only ever accessed with device-scope atomics, the fields above only
     * with plain loads and stores */
    unsigned long long red;
    unsigned int bad;
    char pad2[116];
} ocfd_step_ctl;

/* Parm defaults + Parm::Initialize (Source/Constants.H:9-29, Source/Constants.cpp:7-11).
 * If p->eos_gamma == 0 all defaults are filled first; otherwise only the
 * derived members Rg, cv, cp, kOverMu are recomputed (EBR::read_params,
 * Source/EBR.cpp:522-528). */
int ocfd_parm_initialize(ocfd_parm* p);

/* Build a level from its box list (valid boxes, inclusive lo/hi, 3 ints each):
 * the BoxArray + Geometry + BCRec set-up of EBR (Source/EBR.cpp:43-65,
 * Source/Setup.cpp:17-92,181-205).  Allocates the device work-lists and a
 * dSdt accumulator.  Boxes must not overlap. */
int ocfd_level_create(int nbox, const int* lo, const int* hi, const ocfd_geom* geom, ocfd_level* out);
int ocfd_level_destroy(ocfd_level lev);

/* Number of doubles of a MultiFab on this level with ngrow ghosts and ncomp
 * comps, and the offset (in doubles) of box b's FAB inside it. */
long ocfd_level_mf_size(ocfd_level lev, int ngrow, int ncomp);
long ocfd_level_fab_offset(ocfd_level lev, int box, int ngrow, int ncomp);

/* Cut-cell geometry of the box [lo, hi] grown by ng cells, on the host: the
 * fields EB2::Build provides to the EB box path (Source/Geometry.cpp:9-29,
 * Source/EB/AdvanceBox.cpp:15-31), each an array over the grown box (i
 * fastest): flag (EBCellFlag bit layout: bits 0-1 type 0 regular / 1 cut /
 * 3 covered, bits 2-4 volume count, bit 5+(i+1)+3(j+1)+9(k+1) neighbour
 * (i,j,k) connected), vfrac, apx/apy/apz (area fraction of the cell's low
 * x/y/z face), fcx/fcy/fcz (2 comps: face centroid offsets in cell units),
 * bcent (3 comps: boundary centroid).  Any output pointer may be NULL.
 * Needs no GPU. */
int ocfd_eb_geometry(const ocfd_eb_spec* spec, const ocfd_geom* g, const int* lo, const int* hi, int ng,
                     unsigned* flag, double* vfrac, double* apx, double* apy, double* apz,
                     double* fcx, double* fcy, double* fcz, double* bcent);

/* ocfd_level_create with an embedded boundary: boxes are classified as in
 * EBR::compute_dSdt (Source/Advance.cpp:213,231): covered (every valid cell
 * covered: dSdt = 0), regular (grow(bx,4) all regular: WENO path) or cut
 * (the EB path of Source/EB/AdvanceBox.cpp).  spec NULL or kind
 * OCFD_EB_NONE gives an all-regular level. */
int ocfd_level_create_eb(int nbox, const int* lo, const int* hi, const ocfd_geom* g, const ocfd_eb_spec* spec,
                         ocfd_level* out);

/* Box classification of the level: types[b] = 0 regular, 1 cut, 2 covered. */
int ocfd_level_box_types(ocfd_level lev, int* types);

/* The library's own device tables and scratch (levels, interface registers,
 * multispecies scratch) come from a per-device caching pool: freed blocks are
 * kept and handed out again, since an AMR regrid or a serving loop rebuilds
 * levels all the time (the role of AMReX's caching arena under the reference's
 * EBR levels, Source/EBR.cpp:43-65).  ocfd_cached_device_bytes: bytes held
 * free in the pool; ocfd_release_cached_memory: give them back to HIP (waits
 * for nothing the library still uses). */
long ocfd_cached_device_bytes(void);
int ocfd_release_cached_memory(void);

/* Device memory held by the level itself (work-lists, primitive and partial
 * divergence scratch), in bytes; the caller's MultiFabs are not included.
 * For admission control by HBM footprint. */
long ocfd_level_device_bytes(ocfd_level lev);
int  ocfd_level_nbox(ocfd_level lev);

/* EBR::compute_dSdt regular-box path (Source/Advance.cpp:182-418 with the
 * kernels of the Source/kernel headers): dSdt (0 ghosts, 5 comps) = -div F(S) for every
 * box of the level, S holding 4 filled ghost cells.  dt only feeds flux
 * registers in the reference and is unused here.  flags: OCFD_VISC |
 * OCFD_GRAVITY | OCFD_FLUX_FP32 | OCFD_PLANAR.  stream: hipStream_t (NULL =
 * default stream). */
int ocfd_compute_dsdt(ocfd_level lev, const double* S, double* dSdt, double dt,
                      const ocfd_parm* parm, int flags, void* stream);

/* One RK stage fused with compute_dSdt: S_new(valid) = (c_cur*S_cur +
 * c_old*S_old) + c_dt * dSdt(S_cur), i.e. MultiFab::LinComb + Saxpy of
 * EBR::flow_advance (Source/Advance.cpp:117-148).  S_old may be NULL (then the
 * c_old term is absent).  Ghost cells of S_new are not written. */
int ocfd_rk_stage(ocfd_level lev, const double* S_cur, const double* S_old, double* S_new,
                  double c_cur, double c_old, double c_dt,
                  const ocfd_parm* parm, int flags, void* stream);

/* FillPatch on one level (Source/Advance.cpp:119,126,134 + Exec/<case>/prob_bcfill.cpp):
 * fills the 4 ghost cells of every FAB of S from the level's valid cells,
 * periodic images and the physical BCs at time `time`. */
int ocfd_fill_boundary(ocfd_level lev, double* S, double time, void* stream);

/* Planar fill depth (no counterpart in the reference, whose 2D runs carry all
 * four z ghost layers): with on != 0 the level's ghost fills (this level's
 * FillPatch and the coarse/fine fill of an AMR level) write one z ghost layer
 * instead of four, which is all a z-uniform level stepped with OCFD_PLANAR
 * reads.  Stages without OCFD_PLANAR and the multispecies entry points then
 * return OCFD_EINVAL; levels with cut boxes return OCFD_EUNSUPPORTED. */
int ocfd_level_set_planar(ocfd_level lev, int on);

/* Tile shape of the y and z march sweeps: cells per march chunk (shorter
 * chunks give more independent waves at the cost of a longer halo prologue
 * per cell).  A level starts with 32, halved down to 2 while it has fewer
 * than 2048 march tiles.  Results do not depend on it.  Refused
 * (OCFD_EINVAL) while step graphs of the level exist: a graph holds the tile
 * lists' addresses and launch sizes it was captured with. */
int ocfd_level_set_march_chunk(ocfd_level lev, int ch);
int ocfd_level_march_chunk(ocfd_level lev);

/* Autotuner of the march tile shape: times the y and z sweeps of this level
 * on state S (4 filled ghosts) for chunks of 2, 4, 8, 16, 32 and 64 cells
 * (reps launches each; ms_out[6] gets the mean ms per candidate when not
 * NULL), keeps the fastest and returns it in *ch_out.  With OCFD_PLANAR on a
 * level with a one-layer fill it times the first-layer y march that such a
 * stage runs.  Overwrites the level's scratch only.  Synchronises the stream.
 * Refused while step graphs of the level exist. */
int ocfd_level_autotune(ocfd_level lev, const double* S, const ocfd_parm* parm, int flags, int reps, int* ch_out,
                        double* ms_out, void* stream);

/* EBR::estTimeStep (Source/Advance.cpp:61-79): cfl * min over valid cells of
 * dx_d/(|u_d|+c).  Synchronises the stream and writes the result to *dt_out.
 * Returns OCFD_ENONFINITE (and still writes *dt_out) when a valid cell of S
 * holds a NaN or inf, which the reference's min would pass over silently. */
int ocfd_estdt(ocfd_level lev, const double* S, double cfl, const ocfd_parm* parm,
               double* dt_out, void* stream);

/* One coarse time step of EBR::advance without chemistry (Source/Advance.cpp:81-150):
 * scheme 0 Euler, 1 RK2 (Heun), 2 SSP-RK3.  A holds U^n with filled ghosts on
 * entry and U^(n+1) with filled ghosts on return; B and C are scratch
 * MultiFabs of the same layout (4 ghosts, 5 comps). */
int ocfd_advance(ocfd_level lev, double* A, double* B, double* C, double time, double dt,
                 int scheme, const ocfd_parm* parm, int flags, void* stream);

/* Case initial data into the valid cells of state MultiFab S (4 ghosts, 5 comps):
 * kind 0 = Sod (Exec/0_SOD/prob.H:11-38), 1 = double Mach reflection
 * (Exec/1_DMR/prob.H:11-45), 2 = 3D viscous Kelvin-Helmholtz shear layer
 * (the 3D counterpart of Exec/8_KH/prob.H:11-42, seeded), 3 = uniform
 * inflow over an embedded body (Exec/3_FSTEP/prob.H:11-40: the inflow state,
 * covered cells at rest with the same energy; needs an EB level).  The
 * amrex_probinit/initData pair of the reference (Source/EBR.cpp:111-170).
 * Ghost cells are zeroed.  Synchronous. */
int ocfd_init_state(ocfd_level lev, int kind, double* S, const ocfd_parm* parm, unsigned long long seed);

/* Inflow state of the EB cases (Exec/3_FSTEP/prob.cpp:24-29, literal 0.4) for
 * p, rho, u (defaults of Exec/3_FSTEP/prob_parm.H when the pointer args are
 * negative). */
int ocfd_inflow_state(double p, double rho, double u, double* state);

/* DMR inflow states used by probtype 1 (Exec/1_DMR/prob.cpp:28-40). */
int ocfd_dmr_states(double* left, double* right);

/* Per-kernel timing of one RK stage (the reference's BL_PROFILE scopes,
 * Source/Advance.cpp:184): launches each kernel of ocfd_rk_stage `reps`
 * times between HIP events on `stream` and writes the average duration in ms
 * of the x sweep, y sweep, z sweep (with the stage update), the ghost fill
 * and the primitive recovery to ms_out[0..4].  The ghost fill runs on S_cur.
 * Not for OCFD_PLANAR. */
int ocfd_profile_stage(ocfd_level lev, const double* S_cur, const double* S_old, double* S_new,
                       double c_cur, double c_old, double c_dt, const ocfd_parm* parm, int flags,
                       int reps, double* ms_out, void* stream);

/* ---------------------------------------------------------------------------
 * Chemistry: Li-Dryer 9-species / 19-reaction H2/air mechanism, species order
 * H2 O2 H2O H O OH HO2 H2O2 N2 (Source/Constants.cpp:13-22).
 * Species MultiFabs use the level layout with 4 ghosts and 9 components
 * (Spec_Type, Source/Setup.cpp:242-246); size ocfd_level_mf_size(lev, 4, 9).
 * ------------------------------------------------------------------------- */

/* EBR::chemical_advance (Source/CHEM/ChemAdvance.cpp:7-85): one
 * linearised-implicit step of dt per valid cell (Newton temperature from the
 * sensible energy, production rates and mass-action Jacobian, 9x9 pivoted
 * Gauss solve, clip at 0, heat release into E), in place on S and Spec. */
int ocfd_chem_advance(ocfd_level lev, double* S, double* Spec, double dt, const ocfd_parm* parm, void* stream);

/* Species clip and rescale to rho (EBR::post_timestep, Source/EBR.cpp:333-357). */
int ocfd_species_renorm(ocfd_level lev, const double* S, double* Spec, void* stream);

/* Batched constant-volume reactors (0D ignition).  state: device array of
 * 11 rows of n doubles (SoA): rho, sensible energy density E (J/m^3), then the
 * 9 partial densities (kg/m^3).  ocfd_reactor_init fills it for H2/air at
 * (T0 [K], p0 [Pa], equivalence ratio phi) with X(H2):X(O2):X(N2) = 2phi:1:3.76
 * (CKXTY/CKRHOY/CKUBMS, Source/CHEM/LiDryer.H:21-66,697-715). */
#define OCFD_REACTOR_NCOMP 11
int ocfd_reactor_init(long n, const double* T0, const double* p0, const double* phi, double* state, void* stream);

/* nsteps linearised-implicit steps of dt per reactor: the reference's
 * fixed-step chemistry (Source/CHEM/ChemAdvance.cpp:48-81) applied to a
 * reactor at rest. */
int ocfd_reactor_advance_fixed(long n, double* state, double dt, int nsteps, void* stream);

/* Integration to t_end with per-reactor adaptive substeps: step doubling on
 * the same linearised-implicit step, error max_k |y2-y1|/(atol rho + rtol|y2|)
 * over the partial densities, accepted value 2 y2 - y1 (Richardson, second
 * order), h <- h*clamp(0.9/sqrt(err), 0.2, 4).  One
 * persistent launch; lanes take reactors from a device work counter.
 * steps_out / rejects_out (device int[n], may be NULL): accepted steps (-1 if
 * max_steps was hit) and rejected steps.  work: 8 bytes of device memory. */
int ocfd_reactor_advance_adaptive(long n, double* state, double t_end, double h0, double rtol, double atol,
                                  int max_steps, int* steps_out, int* rejects_out, void* work, void* stream);

/* The same integration in resumable pieces (no counterpart in the reference):
 * begin sets every reactor's time to 0 and step to h0 (th: device double[2n],
 * times then steps) and zeroes the counts; each chunk lets every reactor
 * take at most `budget` substep attempts and stores into *remaining (device
 * unsigned) how many reactors have not reached t_end.  Chunks in a row give
 * results identical to ocfd_reactor_advance_adaptive; a served batch runs in
 * short launches that do not hold a hardware queue for its whole integration. */
int ocfd_reactor_adaptive_begin(long n, double* th, double h0, int* steps, int* rejects, void* stream);
int ocfd_reactor_adaptive_chunk(long n, double* state, double* th, double t_end, double rtol, double atol,
                                int max_steps, int budget, int* steps, int* rejects, unsigned int* remaining,
                                void* stream);

/* Integration to t_end by a Rosenbrock method (no counterpart in the
 * reference, whose chemistry takes fixed linearised-implicit steps,
 * Source/CHEM/ChemAdvance.cpp:48-81): RODAS3 (4 stages, order 3 with an
 * embedded order-2 solution, L-stable; Sandu et al., Atmos. Environ. 31
 * (1997) 3459) on d rho_k/dt = W_k wdot_k with T from the sensible energy,
 * one 9x9 LU per step of the full Jacobian (complete fixed-T derivative of
 * the rates, third bodies and fall-off included, plus the temperature
 * coupling).  Error max_k |K4_k|/(atol rho + rtol max(|y_k|,|y1_k|)); h <- h
 * clamp(0.9 err^(-1/3), 0.2, 6), not growing right after a rejection.
 * Arguments as ocfd_reactor_advance_adaptive (work: 8 bytes). */
int ocfd_reactor_advance_rosenbrock(long n, double* state, double t_end, double h0, double rtol, double atol,
                                    int max_steps, int* steps_out, int* rejects_out, void* work, void* stream);

/* Resumable pieces of the Rosenbrock integration, as the adaptive ones
 * (no counterpart in the reference): th is device double[3n] (times, steps,
 * temperature guesses); chunks in a row give the results of
 * ocfd_reactor_advance_rosenbrock. */
int ocfd_reactor_rosenbrock_begin(long n, double* th, double h0, int* steps, int* rejects, void* stream);
int ocfd_reactor_rosenbrock_chunk(long n, double* state, double* th, double t_end, double rtol, double atol,
                                  int max_steps, int budget, int* steps, int* rejects, unsigned int* remaining,
                                  void* stream);

/* Temperature of each reactor (GET_T_GIVEN_EY, Source/CHEM/LiDryer.H:836-888). */
int ocfd_reactor_temperature(long n, const double* state, double* T, void* stream);

/* ---------------------------------------------------------------------------
 * Multispecies flow (the CHEM build of the reference): state S (5 comps) and
 * partial densities Spec (9 comps), both 4 ghosts.  FP64.  Cut boxes take the
 * EB-CHEM path of EBR::ebchem_compute_dSdt_box (Source/CHEM/FlowAdvanceBox.cpp:
 * species only; the reference leaves their flow dSdt at 0), covered boxes get
 * dSdt = 0.  flags: OCFD_VISC.
 * ------------------------------------------------------------------------- */
#define OCFD_CASE_CHEMSOD 4   /* Exec/5_CHEMSOD: H2/O2 shock tube with a hot driver */

/* FillPatch of the species MultiFab on one level: scalar BCs for every
 * component (Source/Setup.cpp:241-247). */
int ocfd_fill_boundary_spec(ocfd_level lev, double* Spec, double time, void* stream);

/* Species on ext_dir (inflow) sides: ocfd_fill_boundary_spec and the device
 * steps write `spec` (9 partial densities, host array; NULL clears) into
 * every ext_dir ghost cell of the level's species MultiFabs.  Without it those
 * ghost cells keep whatever they hold -- the reference's spec_bcfill leaves
 * them alone too (Exec/9_CHEMEB/prob_bcfill.cpp:37-44), which is undefined for
 * the scratch MultiFabs of its RK stages.  Synchronous. */
int ocfd_level_set_inflow_species(ocfd_level lev, const double* spec);

/* EBR::compute_dSdt_multi, regular boxes (Source/CHEM/FlowAdvance.cpp:92-496):
 * c2prim_rgas, mixture transport (viscous), Steger-Warming split of flow and
 * species with the cell's gamma, WENO-Z5 for 5 + 9 comps, Fickian diffusion,
 * multispecies viscous flux, divop.  dSdt: 5 comps, dSdt_spec: 9 comps, valid
 * layout (ocfd_level_mf_size(lev, 0, 5 or 9)). */
int ocfd_compute_dsdt_multi(ocfd_level lev, const double* S, const double* Spec, double* dSdt, double* dSdt_spec,
                            double dt, const ocfd_parm* parm, int flags, void* stream);

/* No counterpart in the reference: ocfd_compute_dsdt_multi for callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results; the scratch
 * the multispecies path otherwise allocates on its first call (19-component
 * primitives/transport and three 14-component face-flux buffers, 488 B per
 * grown cell of the level) lives in a 64 MiB static device array, held by
 * each GPU the library runs on until the process ends.  A level needing more
 * (more than 137,518 grown cells) is refused with OCFD_EINVAL.  Calls no
 * hipMalloc/hipFree.  Every call on one GPU uses the same array: calls that
 * overlap on one GPU (different streams, concurrently) race. */
int ocfd_compute_dsdt_multi_noalloc(ocfd_level lev, const double* S, const double* Spec, double* dSdt,
                                    double* dSdt_spec, double dt, const ocfd_parm* parm, int flags, void* stream);

/* One stage of EBR::flow_advance_multi for state and species together
 * (LinComb / Saxpy of Source/CHEM/FlowAdvance.cpp:59-86). */
int ocfd_rk_stage_multi(ocfd_level lev, const double* S_cur, const double* Spec_cur, const double* S_old,
                        const double* Spec_old, double* S_new, double* Spec_new, double c_cur, double c_old,
                        double c_dt, const ocfd_parm* parm, int flags, void* stream);

/* EBR::flow_advance_multi (Source/CHEM/FlowAdvance.cpp:9-90): scheme 0 Euler,
 * 1 RK2 (its RK3 is an error in the reference).  A/Aspec in and out, B/Bspec
 * scratch. */
int ocfd_advance_multi(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, double time, double dt,
                       int scheme, const ocfd_parm* parm, int flags, void* stream);

/* EBR::advance with chemistry (Source/Advance.cpp:152-177): iter = 4/2^level
 * (level <= 2, else 1) chemistry steps of dt/(2 iter), the flow step, iter
 * more chemistry steps; renorm != 0 adds the species renormalisation of
 * EBR::post_timestep (Source/EBR.cpp:333-357). */
int ocfd_advance_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, double time, double dt,
                      int level, int scheme, const ocfd_parm* parm, int flags, int renorm, void* stream);

/* Initial state and species of a multispecies case (kind OCFD_CASE_CHEMSOD:
 * Exec/5_CHEMSOD/prob.H:11-82).  Ghost cells are zeroed.  Synchronous. */
int ocfd_init_species_state(ocfd_level lev, int kind, double* S, double* Spec, const ocfd_parm* parm);

/* A uniform premixed state: every cell of S (5 comps) and Spec (9 partial
 * densities), ghost cells included, set to `state` and `spec` (host arrays).
 * The initial data and constant inflow of the reacting channel past a
 * cylinder (the uniform-premixed counterpart of ebr_initdata/ebr_initspec,
 * Exec/9_CHEMEB/prob.H:11-117); ghost cells on an inflow (ext_dir) side keep
 * these species, which the species ghost fill leaves alone there.
 * Synchronous. */
int ocfd_init_species_uniform(ocfd_level lev, const double* state, const double* spec, double* S, double* Spec);

/* Device-driven steps of a multispecies level (no host round trip, graph
 * capturable): ocfd_ctl_start_chem sets time, stop and step limits and the
 * CHEM build's dt = cfl * 1e-8 (ebr_estdt under CHEM, Source/Advance.cpp:54-57;
 * EBR::computeInitialDt, Source/EBR.cpp:189-229).  ocfd_step_device_chem runs
 * one EBR::advance of ocfd_advance_chem -- chemistry x iter, flow_advance_multi
 * (scheme 0 Euler, 1 RK2), chemistry x iter, then the species renormalisation
 * of EBR::post_timestep when renorm != 0 (Source/Advance.cpp:152-177,
 * Source/EBR.cpp:333-357) -- with dt and time from ctl, then the dt update of
 * EBR::computeNewDt; a non-finite cell ends the run with OCFD_ENONFINITE.
 * Bit-identical to the host loop of ocfd_advance_chem calls. */
int ocfd_ctl_start_chem(ocfd_level lev, const double* A, ocfd_step_ctl* ctl, double time, double cfl,
                        double stop_time, int max_steps, const ocfd_parm* parm, void* stream);
int ocfd_step_device_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, ocfd_step_ctl* ctl,
                          int scheme, int level, const ocfd_parm* parm, int flags, int renorm, void* stream);
/* hipGraph of nsteps such steps (rules of ocfd_graph_capture_steps). */
int ocfd_graph_capture_steps_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec,
                                  ocfd_step_ctl* ctl, int scheme, int level, const ocfd_parm* parm, int flags,
                                  int renorm, int nsteps, void* stream, ocfd_graph* out);

/* Initialise device time control: time, cfl, stop_time, max_steps, and the
 * initial dt = cfl * estdt(A) clipped to stop_time (EBR::computeInitialDt,
 * Source/EBR.cpp:189-229,561-563).  No host synchronisation. */
int ocfd_ctl_start(ocfd_level lev, const double* A, ocfd_step_ctl* ctl, double time, double cfl, double stop_time,
                   int max_steps, const ocfd_parm* parm, void* stream);

/* One time step driven by ctl (the stage sequence of ocfd_advance for the
 * scheme, with dt and time read from ctl), followed by the CFL reduction of
 * the new state and the dt update of EBR::computeNewDt (change_max 1.1,
 * stop_time clip, Source/EBR.cpp:231-306).  On a level of regular boxes the
 * reduction is taken by the sweep that finishes the last stage, as each new
 * cell is written.  Once ctl->done is set every kernel of the step returns at
 * once, so extra steps are no-ops.  A non-finite cell in the new state sets
 * ctl->status = OCFD_ENONFINITE and ctl->done.  No host synchronisation, no
 * allocation. */
int ocfd_step_device(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                     const ocfd_parm* parm, int flags, void* stream);

/* hipGraph of nsteps device-driven time steps of one level, captured on
 * `stream` (which must not be the NULL stream) and replayed with
 * ocfd_graph_launch on any stream.  The graph holds the level's work-list
 * addresses and launch sizes: while it exists the level refuses
 * ocfd_level_set_march_chunk, ocfd_level_autotune and ocfd_level_destroy
 * (OCFD_EINVAL); destroy the graphs first. */
int ocfd_graph_capture_steps(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                             const ocfd_parm* parm, int flags, int nsteps, void* stream, ocfd_graph* out);
int ocfd_graph_launch(ocfd_graph g, void* stream);
int ocfd_graph_destroy(ocfd_graph g);

/* ---------------------------------------------------------------------------
 * AMR hierarchies of refinement ratio 2 (regular levels).  Each fine level's
 * domain is the coarser domain refined by 2, its boxes are aligned to 2 and
 * nested in the coarser level.  The subcycled time step that strings these
 * together is the host driver's (Amr::timeStep semantics: a level's step,
 * then two steps of dt/2 of the next finer level, then reflux and
 * average-down; Source/EBR.cpp:231-330).
 * ------------------------------------------------------------------------- */

/* tag_dengrad (Source/Tagging.H:8-36) as called by EBR::errorEst
 * (Source/Tagging.cpp:48-80): tags[cell] = 1 where the larger one-sided
 * density jump >= threshold.  S has filled ghost cells; tags holds one byte per
 * valid cell in the level's 0-ghost layout (box order, i fastest). */
int ocfd_tag_dengrad(ocfd_level lev, const double* S, double threshold, unsigned char* tags, void* stream);

/* ocfd_tag_dengrad with the proper nesting region of the level folded in:
 * flags (one byte per valid cell, the same layout) hold the tag in bit 0 and,
 * in bit 1, whether every cell within Chebyshev distance n_proper of the cell
 * (periodic images included; beyond a non-periodic domain side nothing is
 * required) lies in the level's boxes: the region where the next finer level
 * may place grids (amr.n_proper; Amr::grid_places intersects the tags and the
 * new grids with it).  n_proper 0: bit 1 is set everywhere. */
int ocfd_tag_dengrad_nest(ocfd_level lev, const double* S, double threshold, int n_proper, unsigned char* flags,
                          void* stream);

/* New fine boxes from coarse tags (the AmrLevel regrid's grid_places that
 * errorEst feeds; amr.blocking_factor / amr.grid_eff / amr.max_grid_size of
 * Exec/<case>/inputs): tags dense over [lo, hi] (i fastest, host memory),
 * coarsened by `block` cells, Berger-Rigoutsos clustering to efficiency
 * grid_eff, intersected with the region `mask` allows (dense like tags; the
 * proper nesting domain; NULL = everywhere), chopped to max_size_cells.
 * Boxes (lo[3], hi[3]) in the tag index space go to boxes_out[6*cap];
 * *nbox_out = count.  Host only. */
int ocfd_cluster_tags(const int lo[3], const int hi[3], const unsigned char* tags, const unsigned char* mask, int block,
                      double grid_eff, int max_size_cells, int cap, int* boxes_out, int* nbox_out);

/* The same clustering from tags held in a level's own layout instead of dense
 * domain arrays (a level covers a few percent of its domain): flags, one byte
 * per valid cell of the nbox boxes (lo/hi triples) in the 0-ghost layout
 * (box order, i fastest), as ocfd_tag_dengrad_nest writes them; the cells of
 * the nextra boxes elo/ehi count as tagged too (the projection of the next
 * finer level's new grids in Amr::grid_places).  A cell is tagged when bit 0
 * (or an extra box) and bit 1 hold; with use_mask the clusters are cut back to
 * the cells with bit 1 (cells outside the level's boxes are outside that
 * region).  Returns the boxes ocfd_cluster_tags returns for the equivalent
 * dense tags ((tag OR extra) AND region) and mask (region).  Host only. */
int ocfd_cluster_box_tags(int nbox, const int* lo, const int* hi, const unsigned char* flags, int nextra,
                          const int* elo, const int* ehi, int use_mask, int block, double grid_eff,
                          int max_size_cells, int cap, int* boxes_out, int* nbox_out);

/* Boxes of level f+1 (lo/hi, inclusive, 3 ints each) projected to level f-1: coarsened to level f, grown by
 * n_proper cells there, the parts beyond level f's domain (dom_lo..dom_hi) wrapped to their periodic
 * images on periodic sides and clipped elsewhere, coarsened again; the tags that keep level f's new grids
 * around the finer new grids (Amr::grid_places, AMReX; the reference's regrids run under it,
 * Source/EBR.cpp:231-306).  The boxes may overlap.  Output and capacity as ocfd_cluster_tags.  Host only. */
int ocfd_project_boxes(int nbox, const int* lo, const int* hi, const int dom_lo[3], const int dom_hi[3],
                       const int periodic[3], int n_proper, int cap, int* boxes_out, int* nbox_out);

/* A dense 0/1 mask over a level's domain (nz*ny*nx bytes, i fastest, host
 * memory) grown in place by n cells in every direction (Chebyshev: the tag
 * buffer amr.n_error_buf), or with shrink != 0 shrunk by n cells (the proper
 * nesting domain of a level's coverage, amr.n_proper).  Periodic axes wrap;
 * outside the domain elsewhere counts as unset when growing and set when
 * shrinking.  Host only. */
int ocfd_mask_grow(unsigned char* m, int nx, int ny, int nz, int n, const int* periodic, int shrink);

/* out[0..n) = a*x + b*y: the two-time data combination of FillPatch at a time
 * between a level's old and new states (a = (t1-t)/(t1-t0), b = (t-t0)/(t1-t0)). */
int ocfd_lincomb(long n, double a, const double* x, double b, const double* y, double* out, void* stream);

/* Valid cells of level dst copied from level src (same domain) where src covers them. */
int ocfd_copy_level(ocfd_level dst, double* S_dst, ocfd_level src, const double* S_src, void* stream);

/* Flux register of one coarse/fine interface (EBFluxRegister of EBR,
 * Source/EBR.H flux_reg; CrseAdd/FineAdd at Source/Advance.cpp:344-360,
 * Reflux at Source/EBR.cpp:322-330). */
typedef struct ocfd_cf_s* ocfd_cf;
int ocfd_cf_create(ocfd_level coarse, ocfd_level fine, ocfd_cf* out);
int ocfd_cf_destroy(ocfd_cf c);
int ocfd_cf_nfaces(ocfd_cf c);
/* interface faces as (coarse box, i, j, k, dir, high-face) sextuples */
int ocfd_cf_faces(ocfd_cf c, int* out);
/* zero both registers (flux_reg.reset(), Source/Advance.cpp:98-103) */
int ocfd_cf_reset(ocfd_cf c, void* stream);
/* After an RK stage of the coarse (fine_side 0) or fine (fine_side 1) level:
 * add w_dt * F to every interface face, F recomputed from that stage's
 * primitives (w_dt = the stage's dt weight: RK3 dt/6, dt/6, 2dt/3; RK2 dt/2;
 * Euler dt).  flags as for the stage. */
int ocfd_cf_add(ocfd_cf c, int fine_side, double w_dt, const ocfd_parm* parm, int flags, void* stream);
/* EBR::compute_dSdt with its flux-register arguments (Source/Advance.cpp:182,
 * the CrseAdd/FineAdd block at 344-360; declared at Source/EBR.H:167-169):
 * ocfd_compute_dsdt, then w_dt = dt times this level's interface fluxes added
 * to `fine` (the register joining lev to the next finer level: lev is its
 * coarse side, CrseAdd) and to `current` (the register joining lev to the next
 * coarser level: lev is its fine side, FineAdd).  Either may be NULL.  dt is
 * the stage weight the reference passes (Euler dt, RK2 dt/2, RK3 dt/6, dt/6,
 * 2dt/3).  OCFD_EINVAL when a register does not border lev on that side. */
int ocfd_compute_dsdt_fr(ocfd_level lev, const double* S, double* dSdt, double dt, ocfd_cf fine, ocfd_cf current,
                         const ocfd_parm* parm, int flags, void* stream);
/* copy the registers to host: regc[5*nface], regf[20*nface] (4 sub-faces per face) */
int ocfd_cf_registers(ocfd_cf c, double* regc, double* regf);
/* S_coarse += the flux mismatch of every interface face on the coarse side */
int ocfd_cf_reflux(ocfd_cf c, double* S_coarse, void* stream);
/* The coarse part of the two-level FillPatch (FillPatchTwoLevels with the
 * state's cell-conservative linear interpolater, Source/Setup.cpp:196): after
 * ocfd_fill_boundary(fine, S_fine, t), every fine ghost cell whose source cell
 * no fine box owns is interpolated from S_coarse, the coarse state at time t
 * with filled ghost cells. */
int ocfd_cf_fill_fine(ocfd_cf c, double* S_fine, const double* S_coarse, void* stream);

/* Valid cells of a new fine level interpolated from the coarse state (the
 * same interpolater; AmrLevel::init of a regridded level), skipping cells the
 * old fine level covers when old_fine is given. */
int ocfd_cf_interp_fine(ocfd_cf c, double* S_fine, const double* S_coarse, ocfd_level old_fine, void* stream);

/* EBR::avgDown (Source/EBR.cpp:535-559): coarse cells under the fine level
 * take the mean of their 8 fine cells; ncomp components (5 state, 9 species). */
int ocfd_cf_avg_down(ocfd_cf c, const double* S_fine, double* S_coarse, int ncomp, void* stream);

/* One level of an AMR hierarchy during a time step (MultiFabs with 4 ghosts,
 * 5 comps; the caller keeps the array across calls). */
typedef struct ocfd_amr_level_state {
    double* old;     /* the level's state at t_old */
    double* cur;     /* the level's state at t_new (the current state) */
    double* s1;      /* stage scratch */
    double* s2;      /* stage scratch (SSP-RK3) */
    double* tmp;     /* the state at an intermediate time (used when a finer level fills from it) */
    double t_old, t_new;
} ocfd_amr_level_state;

/* Advance level l of an nlev hierarchy from t by dt (AmrLevel::advance inside
 * Amr::timeStep, EBR::flow_advance of Source/Advance.cpp:81-150): st[l].old and
 * st[l].cur are swapped (cur then holds the state at t) and the stages write the
 * new state into st[l].cur; each stage's ghost cells come from the level and,
 * recursively, from every coarser level m at the stage time -- its old/new pair
 * combined between st[m].t_old and st[m].t_new and ghost-filled the same way
 * (FillPatch with the two-level interpolater).  Register adds after each stage:
 * cfs[l+1] as the coarse side (reset first), cfs[l] as the fine side.
 * cfs[m] joins levels m-1 and m (cfs[0] unused).  The caller runs the finer
 * levels' subcycles after this call and then ocfd_cf_reflux / ocfd_cf_avg_down
 * (EBR::post_timestep, Source/EBR.cpp:313-330).  No host synchronisation. */
int ocfd_amr_advance_level(int nlev, const ocfd_level* levels, const ocfd_cf* cfs, ocfd_amr_level_state* st, int l,
                           double t, double dt, int scheme, const ocfd_parm* parm, int flags, int do_reflux,
                           void* stream);

/* EBR::estTimeStep of every level of a hierarchy in one call, the loop of
 * Amr::computeNewDt over the levels (Source/Advance.cpp:61-79): dt_out[l] =
 * cfl * min over level l's valid cells of dx_d/(|u_d|+c), for states[l] of
 * levels[l].  All reductions are queued, then the results come back in one
 * copy and one synchronisation of the stream.  Returns OCFD_ENONFINITE (all of
 * dt_out still written) when a level's state holds a NaN or inf. */
int ocfd_amr_estdt(int nlev, const ocfd_level* levels, const double* const* states, double cfl,
                   const ocfd_parm* parm, double* dt_out, void* stream);

/* Ghost cells of X, a state of level l at time t, filled as a stage's are in
 * ocfd_amr_advance_level (FillPatch of level l: the level, then the coarser
 * levels at time t, recursively).  For regrids and output. */
int ocfd_amr_fill_level(int nlev, const ocfd_level* levels, const ocfd_cf* cfs, ocfd_amr_level_state* st, int l,
                        double* X, double t, void* stream);

/* ---------------------------------------------------------------------------
 * On-disk formats
 * ------------------------------------------------------------------------- */

/* Plot variables: rho xmom ymom zmom Eden (State_Type, Source/Setup.cpp:202-206),
 * Cost (Cost_Type, written as 0), ux uy uz and pressure (the derived
 * "velocity" and "pressure" of derive_plot_vars, Source/Setup.cpp:94-179,
 * Exec/1_DMR/inputs) and vfrac (the EB volume fraction AmrLevel adds). */
#define OCFD_PLOT_NVARS 11

/* The plot variables of every valid cell of S (4 ghosts, 5 comps) into out
 * (device, the level's 0-ghost layout with OCFD_PLOT_NVARS comps). */
int ocfd_plot_vars(ocfd_level lev, const double* S, double* out, const ocfd_parm* parm, void* stream);

/* Write an AMReX "HyperCLaw-V1.1" plotfile (EBR::thePlotFileType,
 * Source/EBR.H:33-35; the AmrLevel::writePlotFile layout of Source/IO.cpp:32-262):
 * dir/Header, dir/Level_<l>/Cell_H and dir/Level_<l>/Cell_D_00000 for nlev
 * levels (each refining the one below by 2), S[l] the state of level l,
 * level_steps[l] its step count.  Synchronises the stream. */
int ocfd_plotfile_write(const char* dir, int nlev, const ocfd_level* levels, const double* const* S, double time,
                        const int* level_steps, const ocfd_parm* parm, void* stream);

/* fcompare (Util/Plotfile/fcompare.cpp:19-390): compare two plotfiles level by
 * level and variable by variable in norm 0 (max), 1 or 2 (the latter scaled by
 * the cell volume to the power 1/norm), relative to plotfile A.  *result: 1
 * identical, 2 within the tolerances (absolute atol or relative rtol per
 * variable), 0 different, -1 NaN present.  The fcompare report text goes to
 * report[report_cap].  Host only; needs no GPU. */
int ocfd_plotfile_compare(const char* file_a, const char* file_b, int norm, double rtol, double atol,
                          int allow_diff_grids, int* result, char* report, long report_cap);

/* Checkpoints (AmrLevel::checkPoint / restart, Source/IO.cpp:9-30): the valid
 * cells of a level's MultiFab S (4 ghosts, ncomp comps) in VisMF form,
 * dir/<name>_H and dir/<name>_D_00000 (the FAB format of the plotfiles), and
 * back into S (ghost cells zeroed) for a level with the same boxes.  The
 * hierarchy's times, steps, dts and grids go in a text header the driver
 * writes beside them.  Synchronise the stream. */
int ocfd_checkpoint_write_level(const char* dir, const char* name, ocfd_level lev, const double* S, int ncomp,
                                void* stream);
int ocfd_checkpoint_read_level(const char* dir, const char* name, ocfd_level lev, double* S, int ncomp, void* stream);

/* Last error message of the calling thread ("" if none). */
const char* ocfd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OCFD_H */
