/* hfm.h — C ABI of libhfm.so, the MI355X (gfx950) air/heat substep of Heatflow-MPM.
 *
 * The reference exposes this path as the Python class heatflow_mpm.solver.HeatflowMPM
 * (heatflow_mpm/solver.py:41-110) on top of AirGrid (heatflow_mpm/grid.py:177-263) and
 * HeatField (heatflow_mpm/heat.py:177-210).  Every entry point below replaces one piece of
 * that class; the Python package heatflow_mpm (heatflow-mpm_amd/heatflow_mpm) binds them with
 * ctypes and keeps the reference's method names.
 *
 * Conventions
 *   - Every function returns 0 (HFM_OK) or a negative HFM_ERR_* code; hfm_last_error() gives
 *     the message of the last failure on the calling thread.
 *   - Host buffers are owned by the caller and are copied; device buffers are owned by the
 *     scene.  All arrays are C-order with the z index fastest, like the reference's Warp arrays:
 *       cell lattice   (n,   n,   n  )  flat index (i*n + j)*n + k
 *       x-face lattice (n+1, n,   n  )  flat index (i*n + j)*n + k
 *       y-face lattice (n,   n+1, n  )  flat index (i*(n+1) + j)*n + k
 *       z-face lattice (n,   n,   n+1)  flat index (i*n + j)*(n+1) + k
 *   - One host thread per scene.  Each scene owns one HIP stream (default priority); different
 *     scenes may be driven from different threads concurrently.
 *   - CG non-convergence is reported in hfm_stats, not as an error (the reference discards
 *     the residual: grid.py:248, heat.py:209).
 */
#ifndef HFM_H
#define HFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFM_OK 0
#define HFM_ERR_ARG (-1)       /* bad argument (null pointer, size, field id) */
#define HFM_ERR_HIP (-2)       /* a HIP runtime call failed */
#define HFM_ERR_OOM (-3)       /* device allocation failed */
#define HFM_ERR_NONFINITE (-4) /* a non-finite value was detected in inputs or state */
#define HFM_ERR_STATE (-5)     /* call not valid in the current scene state */

/* Scene parameters.  Field names follow GridConfig (grid.py:13-22), Material (mpm.py:9-22),
 * the module constants of solver.py:15-16 and the hard-coded solve settings of
 * grid.py:248 / heat.py:198-209. */
typedef struct hfm_params {
    int32_t n;                /* cells per side, int(round(L/h)) (grid.py:183) */
    int32_t buoyancy;         /* GridConfig.buoyancy (grid.py:19) */
    int32_t cg_max_iter;      /* GridConfig.cg_max_iter, pressure PCG (grid.py:22) */
    int32_t cg_check_every;   /* 20 (grid.py:248) */
    int32_t heat_max_iter;    /* 200 (heat.py:198) */
    int32_t heat_check_every; /* 10 (heat.py:209) */
    int32_t precision;        /* HFM_PREC_* */
    int32_t flags;            /* HFM_FLAG_* */
    double h;                 /* cell size [m] (grid.py:16) */
    double rho;               /* air density (grid.py:17) */
    double t_amb;             /* room temperature [C] (grid.py:18) */
    double gravity;           /* 9.81 (grid.py:219) */
    double alpha_f_min;       /* floor of the air fraction (grid.py:20) */
    double cg_tol;            /* pressure PCG relative tolerance (grid.py:21) */
    double heat_tol;          /* heat PCG relative tolerance, 1e-6 (heat.py:198) */
    double kappa_air;         /* KAPPA_AIR or the constructor argument (solver.py:15,43) */
    double kappa_mat;         /* Material.kappa (mpm.py:17) */
    double vol;               /* particle volume spacing^3 (solver.py:65) */
    double k_particle;        /* K_PARTICLE (solver.py:16) */
    double t_sol, t_liq;      /* melt window (mpm.py:12-13) */
    double yield_solid;       /* mpm.py:14 */
    double mu_solid;          /* viscosity_solid (mpm.py:15) */
    double mu_liquid;         /* viscosity_liquid (mpm.py:16) */
    double t_max;             /* max(t_rod, t_cap, t_amb, t_init) (solver.py:57) */
    int32_t dir_planes;       /* pressure pass A layout (no counterpart in the reference): 0 chosen by
                                 grid size, 4 / 8 / 16 planes per LDS tile, -1 untiled */
} hfm_params;

#define HFM_PREC_FP32 0       /* the reference's arithmetic: fp32 storage and math */
/* Mixed precision (no counterpart in the reference).  State (T, u, p) stays fp32.  Both PCG
 * solves become iterative refinement: an outer step forms the true residual r = b - A x in fp32
 * arithmetic with its norm in fp64 and stops when ||r|| <= max(tol * ||b||, tol) (the
 * reference's threshold, now on the true residual); otherwise an inner Jacobi-PCG on r scaled to
 * unit RMS, with r, z, p and A p stored as FP16 (fp32 arithmetic, fp64 dot products), reduces it
 * by 1e-2 and adds its correction to x.  Advection traces its departure points through the fp32
 * velocity.  The iteration cap counts inner iterations; counts are
 * reported exactly (no check_every quantisation).  Not combinable with HFM_FLAG_FUSED_CG. */
#define HFM_PREC_MIXED 1

#define HFM_FLAG_TIME_KERNELS 1  /* record HIP events around every CG kernel launch */
/* Pressure PCG in one launch per iteration instead of two (no counterpart in the reference):
 * warp's Jacobi PCG in the single-reduction form of Chronopoulos and Gear (u = D r, w = A u,
 * s = A p by the recurrence s = w + beta s, p.Ap = w.u - beta^2 p.Ap_old), the same iterates in
 * exact arithmetic, one global reduction and one launch per iteration.  Same stopping rule and
 * count; results agree with the default form to rounding, iteration counts within one.  Used where
 * the pressure pass runs tiled (the autotuned layout picks it where it is faster). */
#define HFM_FLAG_FUSED_CG 2
/* Pressure PCG preconditioned by a geometric multigrid V-cycle instead of Jacobi (no counterpart in
 * the reference): cell-centred levels halving n down to at most 8 cells per side (n = m 2^k,
 * m <= 8, n >= 16), face coefficients averaged from the fine faces, restriction by summing the 8
 * children, injection, one weighted-Jacobi sweep before and after the coarse correction, 16 on the
 * coarsest level.  Same stopping rule and iteration as warp's CG, far fewer iterations; results
 * agree with the Jacobi solve to the solve's tolerance, not bit for bit.  With HFM_PREC_MIXED it
 * preconditions the FP16 inner solves of the refinement.  Not with HFM_FLAG_FUSED_CG.  A tolerance
 * below what fp32 reaches (about 1e-7 here) runs each solve to its cap or to a breakdown stop. */
#define HFM_FLAG_MG_PCG 4
/* Limited MacCormack advection of T and of u, v, w (Selle et al., J. Sci. Comput. 35, 2008): the
 * semi-Lagrangian value, corrected by half the error of tracing it back to the start, clamped to
 * the range of the eight values around the departure point.  No counterpart in the reference's
 * substep: it prepares the limiter (value_range, fields.py:51-65) but advects semi-Lagrangian
 * only.  fp32 only (not with HFM_PREC_MIXED). */
#define HFM_FLAG_MACCORMACK 8
/* With HFM_FLAG_FUSED_CG: run every iteration of a pressure solve in one cooperative launch of one
 * workgroup per tile, all resident for the whole solve (no counterpart in the reference).  Each
 * workgroup keeps its tile's vectors on chip and exchanges only the tile's faces and its partial
 * sums with the others through memory; the results, scalars and iteration counts are those of
 * HFM_FLAG_FUSED_CG bit for bit.  Used when the device keeps all of the scene's tiles resident at
 * once (n <= 96 on MI355X: 216 tiles of 16 x 16 x 16 cells, one per CU); in a batch such scenes are
 * packed into cooperative launches of at most the device's resident workgroups, each scene's
 * workgroups waiting only on each other.  A scene whose tiles do not fit runs HFM_FLAG_FUSED_CG as if
 * this flag were absent.  Persistent launches on one device run one at a time.  A workgroup that
 * waits longer than a fixed number of polls for the others ends the solve with HFM_ERR_HIP. */
#define HFM_FLAG_PERSISTENT_CG 16

/* Per-substep statistics of the last hfm_substep. */
typedef struct hfm_stats {
    int32_t p_iters;          /* pressure PCG count as warp.optim.linear.cg returns it
                                 (multiple of check_every, or the iteration cap): AirGrid.cg_iters */
    int32_t p_iters_exact;    /* iterations that changed the pressure */
    int32_t h_iters;          /* heat PCG count as warp returns it: HeatField.cg_iters */
    int32_t h_iters_exact;
    int32_t p_converged;      /* 1 when ||r||^2 <= atol^2 at exit, 0 at the iteration cap; 2: stopped at the
                                 fp32 floor of the residual, above atol (HFM_PREC_MIXED: a refinement step cut
                                 the true residual by less than sqrt 2, or with HFM_FLAG_MG_PCG by less
                                 than 4 after an inner solve that reached its tolerance; HFM_FLAG_MG_PCG in fp32: ||r||^2 set no new
                                 minimum for 8 iterations or grew 100 times above its minimum, where the
                                 multigrid-preconditioned iteration would diverge).  -1: stopped on a breakdown (no counterpart in warp, which
                                 runs on to its cap): p.Ap or r.z not positive, alpha or beta not finite,
                                 or a non-finite ||r||^2.  x keeps its last iterate; if ||r||^2 is not
                                 finite the substep returns HFM_ERR_NONFINITE */
    int32_t h_converged;
    int64_t substeps;         /* substeps run since creation */
    double p_rr;              /* final ||r||^2 of the pressure solve */
    double p_atol;            /* max(tol * ||b||, tol) */
    double h_rr;
    double h_atol;
    double rhs_mean;          /* mean removed from the pressure rhs (grid.py:239) */
    int64_t p_iters_total;    /* sum of p_iters_exact over all substeps */
    int64_t h_iters_total;    /* sum of h_iters_exact over all substeps */
    int32_t p_outer;          /* HFM_PREC_MIXED: refinement steps of the pressure solve (0 in fp32) */
    int32_t h_outer;          /* HFM_PREC_MIXED: refinement steps of the heat solve */
    double source_dT_sum;     /* the heater energy balance of the last substep, no counterpart in the
                                 reference: sum over source cells of the temperature rise the capped source
                                 step gave (heat.py:154-158), fp32 rises summed in fp64 on the device in a
                                 fixed order.  Times rho c_p h^3 it is the energy put into the air, equal to
                                 P_W dt (heater.py:52) while no source cell is at its cap; 0 without a
                                 source */
} hfm_stats;

/* Field identifiers for hfm_get_field / hfm_set_field. */
enum hfm_field {
    HFM_FIELD_T = 0,      /* cell temperature, HeatField.T (heat.py:181) */
    HFM_FIELD_T2,         /* HeatField.T2: after a substep, the previous T (swap, solver.py:85) */
    HFM_FIELD_TW,         /* blended/fixed/sourced T before advection (solver.py:74-77) */
    HFM_FIELD_U,          /* x-face velocity (grid.py:186) */
    HFM_FIELD_V,          /* y-face velocity */
    HFM_FIELD_W,          /* z-face velocity */
    HFM_FIELD_P,          /* cell pressure, warm start of the next solve (grid.py:192,248) */
    HFM_FIELD_RHS,        /* pressure rhs after mean removal (grid.py:237-241) */
    HFM_FIELD_ALPHA_S,    /* static + particle solid fraction (grid.py:192,206-213) */
    HFM_FIELD_ALPHA_P,    /* particle-only solid fraction, HeatField.alpha (heat.py:12-18) */
    HFM_FIELD_TSUM,       /* particle heat sum, HeatField.tsum */
    HFM_FIELD_KAPPA,      /* cell diffusivity (heat.py:44-47) */
    HFM_FIELD_AX, HFM_FIELD_AY, HFM_FIELD_AZ,   /* face air fractions (grid.py:46-59) */
    HFM_FIELD_QX, HFM_FIELD_QY, HFM_FIELD_QZ,   /* particle flux, zeroed on walls (solver.py:30-32) */
    HFM_FIELD_KX, HFM_FIELD_KY, HFM_FIELD_KZ,   /* relaxation weights, particle + static */
    HFM_FIELD_MX, HFM_FIELD_MY, HFM_FIELD_MZ,   /* relaxation momenta, particle + static */
    HFM_FIELD_PDIAG,      /* pressure Jacobi preconditioner (grid.py:118-137) */
    HFM_FIELD_HRHS,       /* heat rhs (heat.py:108-118) */
    HFM_FIELD_HDIAG,      /* heat Jacobi preconditioner (heat.py:121-144) */
    HFM_FIELD_COUNT
};

typedef struct hfm_scene hfm_scene;

/* HeatflowMPM.__init__ allocation part: AirGrid + HeatField state (grid.py:181-196,
 * heat.py:178-187).  T, T2, Tw start at t_amb, velocities and pressure at zero. */
int hfm_scene_create(const hfm_params* params, int device, hfm_scene** out);
int hfm_scene_destroy(hfm_scene* s);

/* AirGrid.set_static (grid.py:198-200): static solid fraction (cell lattice) and static
 * relaxation weights / momenta per face lattice, added every substep (grid.py:206-213). */
int hfm_set_static(hfm_scene* s, const float* alpha, const float* kx, const float* ky,
                   const float* kz, const float* mx, const float* my, const float* mz);

/* HeatField.set_fixed (heat.py:189-192): Dirichlet cells; also applies them to T. */
int hfm_set_fixed(hfm_scene* s, const int32_t* fixed, const float* tfix);

/* HeatField.set_source (heat.py:194-196): capped volumetric source on mask cells. */
int hfm_set_source(hfm_scene* s, const int32_t* mask, double rate, double t_cap);

/* The material enters as particle positions, velocities and temperatures (the state the
 * reference reads from ThermalMPM: mpm.py:54-60 and solver.py:60).  P may be 0. */
int hfm_set_particles(hfm_scene* s, int64_t P, const float* pos, const float* vel,
                      const float* temp);

/* Positions/velocities only (a moving body between substeps); temperatures are kept. */
int hfm_update_particles(hfm_scene* s, const float* pos, const float* vel);

/* HeatflowMPM.substep(dt) without Newton's b.step (solver.py:63-86). */
int hfm_substep(hfm_scene* s, double dt);

/* K substeps back to back; *ms (may be NULL) receives the HIP-event time on the scene's
 * stream from before the first to after the last substep. */
int hfm_run(hfm_scene* s, double dt, int32_t K, double* ms);

/* Copy a field to / from the host (hfm_field; lattice size per the table above). */
int hfm_get_field(hfm_scene* s, int32_t field, float* out);
int hfm_set_field(hfm_scene* s, int32_t field, const float* in);
int64_t hfm_field_size(const hfm_scene* s, int32_t field);

/* AirGrid.cell_velocity (grid.py:156-162,253-255): (n^3, 3) cell-centred velocity. */
int hfm_get_cell_velocity(hfm_scene* s, float* out);

/* Per-particle temperature, melt fraction, yield stress and viscosity (heat.py:161-174). */
int hfm_get_particles(hfm_scene* s, float* temp, float* melt, float* yield_stress,
                      float* viscosity);

/* Debug view of the particle binning, no counterpart in the reference (whose scatter forms the
 * indices inside its atomic loop, fields.py:39-48): for every particle p (in the order given to
 * hfm_set_particles), bin[p] = the cell (i n + j) n + k it is sorted into (floor(pos / h), clamped to
 * the box), and base[12 p + 3 l + d] / frac[12 p + 3 l + d] the index int(floor(g)) and fraction
 * g - floor(g) of g = pos_d / h - off_d on lattice l (0 cells, 1 x faces, 2 y faces, 3 z faces; the
 * offsets of fields.py:84-90), computed on the device from the stored pos / h exactly as the particle
 * gather reads it (fields.py:5-10).  Bins the particles first if they changed since the last
 * substep. */
int hfm_get_particle_bins(hfm_scene* s, int32_t* bin, int32_t* base, float* frac);

/* AirGrid.divergence_max (grid.py:257-260) and AirGrid.speed_max (grid.py:262-263). */
int hfm_diagnostics(hfm_scene* s, double* divergence_max, double* speed_max);

int hfm_get_stats(hfm_scene* s, hfm_stats* out);

/* With kernel timing on (HFM_FLAG_TIME_KERNELS at creation, or hfm_set_kernel_timing): total
 * time and launch count of one kernel class (HFM_KT_*) since the last reset, from HIP events
 * recorded around every launch on the scene's stream. */
#define HFM_KT_PCG_DIR 0      /* pressure PCG pass A: direction update + SpMV + p.Ap */
#define HFM_KT_PCG_UPD 1      /* pressure PCG pass B: x, r update + r.r, r.z */
#define HFM_KT_PCG_FUSED 2    /* pressure PCG, HFM_FLAG_FUSED_CG: one whole iteration */
#define HFM_KT_PCG_PERSIST 3  /* pressure PCG, HFM_FLAG_PERSISTENT_CG: one whole solve */
#define HFM_KT_COUNT 4
int hfm_kernel_time(hfm_scene* s, int32_t kernel_class, double* total_ms, int64_t* launches);
int hfm_kernel_time_reset(hfm_scene* s);

/* Switch the kernel timing of HFM_FLAG_TIME_KERNELS on (1) or off (0) after creation. */
int hfm_set_kernel_timing(hfm_scene* s, int32_t on);

int hfm_synchronize(hfm_scene* s);

/* Device memory the scene holds, in bytes (fields, solver vectors, particle arrays, tables);
 * no counterpart in the reference.  Used to admit scenes against an HBM budget. */
int64_t hfm_scene_bytes(const hfm_scene* s);

/* ---- Batches: many scenes stepped together ------------------------------------------------
 * No counterpart in the reference, which steps one HeatflowMPM at a time (solver.py:63-86):
 * a batch runs the same substep for every scene it holds, each stage one kernel launch over all
 * of them, so small grids fill the GPU together.  Scenes may differ in size, particles, heater
 * and solver settings; each keeps its own PCG solves (own tolerances, iteration counts, stats)
 * and its results are the same as when it is stepped alone.  All scenes must be on one device.
 * A scene may be in several batches and still be stepped alone; work on it is ordered across
 * them.  Destroy a batch before the scenes it holds (hfm_scene_destroy returns HFM_ERR_STATE
 * while a batch holds the scene).  One host thread per batch. */
typedef struct hfm_batch hfm_batch;
int hfm_batch_create(hfm_scene* const* scenes, int32_t count, hfm_batch** out);
int hfm_batch_destroy(hfm_batch* b);
/* Replace the scenes a batch holds (count may be 0), keeping its stream and buffers: for a
 * serving loop whose scenes join and leave between substeps. */
int hfm_batch_set_scenes(hfm_batch* b, hfm_scene* const* scenes, int32_t count);
/* One substep of every scene (same dt). */
int hfm_batch_substep(hfm_batch* b, double dt);
/* K substeps back to back; *ms (may be NULL) receives the HIP-event time of the batch. */
int hfm_batch_run(hfm_batch* b, double dt, int32_t K, double* ms);
int hfm_batch_synchronize(hfm_batch* b);
/* Kernel timing of the batch's launches, as hfm_kernel_time / hfm_set_kernel_timing. */
int hfm_batch_set_kernel_timing(hfm_batch* b, int32_t on);
int hfm_batch_kernel_time(hfm_batch* b, int32_t kernel_class, double* total_ms, int64_t* launches);
int hfm_batch_kernel_time_reset(hfm_batch* b);

/* Number of HIP devices visible (0 without a GPU). */
int hfm_device_count(void);

/* ---- Seeding material in a mesh ------------------------------------------------------------
 * Replaces the Warp kernel `_inside` of shapes.py:13-17 as launched by sample_interior
 * (shapes.py:35-44): flags[i] = 1 where the winding number of the closed triangle mesh at
 * points[i] exceeds `threshold` (the reference uses 0.5), else 0.  tri: nt triangles as 9 floats
 * each (three xyz vertices, counter-clockwise seen from outside); points: np xyz.  The winding
 * number is the exact sum of the triangles' solid angles over 4 pi (Warp's query approximates the
 * same sum); `winding` (may be NULL) receives it per point.  Host buffers in and out; the call
 * copies, runs on `device` and returns when the results are written. */
int hfm_mesh_inside(const float* tri, int64_t nt, const float* points, int64_t np, float threshold, int32_t* flags,
                    float* winding, int device);

/* hfm_mesh_inside without device allocations: the same results (flags and winding numbers bit for bit),
 * with the device copies of the triangles, points and results in a fixed 8 MiB array per GPU, held until
 * the process ends, instead of hipMalloc'd buffers.  No counterpart in the reference; for callers that
 * cannot allocate device memory (loops under a no-allocation rule, harnesses that forbid hipMalloc).
 * Limit: round64(9 nt) + round64(3 np) + 2 round64(np) <= 2^21 words (round64 rounds up to a multiple of
 * 64), e.g. 5 120 triangles and 400 000 points; larger calls return HFM_ERR_ARG.  Calls share the array:
 * calls on one GPU are serialised inside the library, so concurrent host threads wait for each other. */
int hfm_mesh_inside_noalloc(const float* tri, int64_t nt, const float* points, int64_t np, float threshold,
                            int32_t* flags, float* winding, int device);

const char* hfm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HFM_H */
