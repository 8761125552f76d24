/*
 * morphompm_hip.h -- C ABI of the MI355X (gfx950) MLS-MPM growth step and its
 * hand-written adjoint.
 *
 * One shared library (morphompm_amd/_lib/libmorphompm_hip.so) exports these
 * symbols.  Conventions:
 *   - every buffer is a caller-owned device pointer; the library never
 *     allocates device memory (mpm_step_fwd_noalloc_f32 uses a fixed scratch
 *     array the library holds on each GPU instead of a workspace argument);
 *   - calls are stream-ordered and asynchronous on `stream` (a hipStream_t,
 *     passed as void*; NULL is the default stream);
 *   - functions return MPM_OK (0) or a negative MPM_E* code and never throw or
 *     abort across the ABI; mpm_last_error() holds a per-thread message;
 *   - _f32 / _f64 pairs share one templated implementation.
 *
 * Particle layout (structure of arrays, row stride `n_slots`):
 *   dyn  [24][n_slots]: rows 0-2 x, 3-5 v, 6-14 C, 15-23 F (3x3 row-major)
 *   Fg   [9][n_slots] : the growth tensor G; the step uses
 *                       Fg = I + a (G - I) with a = sched[instance]
 *                       (sched == NULL means a = 1, i.e. Fg = G exactly)
 * Grid layout: grid [n_nodes][4]: per node (mass, vx, vy, vz).
 * Sort record for one step: sortbuf int32 [2 n_slots + n_bins + 2 n_tiles + 2]:
 *   rank[n_slots] | sorted[n_slots] | bin_start[n_bins + 1] |
 *   unit_rank[n_tiles] | active_unit[n_tiles] | n_active[1]
 *   Only the units (tiles) that hold particles this step get bins: unit u has
 *   compact index unit_rank[u] (-1 when empty), active_unit inverts it, and
 *   bin_start covers n_active * TB bins.  Each instance's sorted range is its
 *   own slot range; padding slots sort into the instance's dead unit.
 * P2G payload for one step: payload [n_slots][16] in slot order, per
 *   particle (fx[3], q[3], A'[9], cell) -- the forward step's momentum map
 *   m_i += w pmass, mom_i += w (q + A' o) for stencil offset o, and the packed
 *   cell ((base+2) per axis in 10 bits, all ones off the grid).  Grids are
 *   limited to 1022 nodes per axis by this packing (base+2 <= 1023); the
 *   Python plan rejects larger grids before any launch.
 *
 * Reference interfaces these replace (paths in 2nugu/morphompm):
 *   python/morphompm/constitutive.py:1-12 (stress / stress_vjp seam),
 *   python/morphompm/transfer.py:33 (step), :82 (step_vjp).
 */
#ifndef MORPHOMPM_HIP_H
#define MORPHOMPM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 7
#define MPM_SLOT_ALIGN 256 /* particle and node ranges of an instance start at multiples of this */

enum { MPM_OK = 0, MPM_EINVAL = -1, MPM_ELAUNCH = -2, MPM_ENOTSUP = -3 };
enum { MPM_NEO_HOOKEAN = 0, MPM_HENCKY = 1, MPM_HERSCHEL_BULKLEY = 2 };
/* Rows of dyn_out a chained step writes (mpm_step_fwd_chain_*): all of them
 * (a checkpoint or a returned state), all but v (a step record: the reverse
 * pass reads x, C and F of it), or x and F only (a state read by the next
 * chained step alone).  Rows not written keep their previous contents.
 * MPM_OUT_NONE runs no G2P at all: the step writes its record (grid, sort
 * record) and not dyn_out, which may be NULL -- the last step of a
 * recomputed checkpoint segment, whose output state is the next checkpoint
 * already held. */
enum { MPM_OUT_FULL = 0, MPM_OUT_RECORD = 1, MPM_OUT_NEXT = 2, MPM_OUT_NONE = 3 };

/* One independent simulation (python/morphompm/config.py:21-40 SimConfig plus
 * its constitutive model).  Node grid nx*ny*nz covers [0, n*dx) per axis.
 * Particles are binned by cell base+2 in [0, n+2) per axis; with b' = bin +
 * shift, the bins are grouped into tiles ("units") of tile[0]*tile[1]*tile[2]
 * bins (tx*ty*tz of them, tx = ceil((nx+2+shift[0])/tile[0]) ...), numbered
 *   unit = t_off + (itx*ty + ity)*tz + itz,  itx = b'x / tile[0],
 * with the bin's place in its unit from b' mod tile, followed by one dead unit
 * (t_off + tx*ty*tz) that holds particles with no in-grid node.  An instance
 * owns tx*ty*tz + 1 units. */
typedef struct mpm_instance {
  int64_t p_off; /* first particle slot, multiple of MPM_SLOT_ALIGN */
  int64_t g_off; /* first node, multiple of MPM_SLOT_ALIGN */
  int64_t reserved;
  int64_t t_off; /* first unit */
  int32_t n;     /* live particles */
  int32_t nx, ny, nz;
  int32_t tx, ty, tz; /* P2G tiles per axis */
  int32_t model;      /* MPM_NEO_HOOKEAN | MPM_HENCKY | MPM_HERSCHEL_BULKLEY */
  int32_t shift[3];   /* tile-grid offset in bins, 0 <= shift < tile: bin b lies in
                         tile (b + shift) / tile; aligns tiles with occupied cells */
  int32_t pad;
  int32_t origin[3];  /* positions are stored relative to node origin*dx (whole cells),
                         so FP32 keeps more bits of small displacements; node = base + origin */
  int32_t pad2;
  double dx, dt, damping, pmass, vol0;
  double prm[6]; /* NH/Hencky: mu, lambda; HB: kappa, K, n, tau_y, m, eps */
} mpm_instance;

/* A batch of instances laid out for one launch per kernel.  The struct lives
 * in host memory; the pointers in it are device pointers. */
typedef struct mpm_batch {
  int32_t n_inst;
  int32_t tile[3];   /* P2G tile shape in bins (x, y, z) */
  int64_t n_slots;   /* particle slots incl. padding, multiple of MPM_SLOT_ALIGN */
  int64_t n_nodes;   /* node slots incl. padding, multiple of MPM_SLOT_ALIGN */
  int64_t n_bins;    /* bin capacity: n_tiles * tile volume */
  int64_t n_tiles;   /* units of all instances (tiles + one dead unit each) */
  int32_t block;     /* threads per block of the particle kernels: 64, 128 or 256 */
  int32_t model_mask; /* bit m set when some instance uses model m; one bit set selects
                         kernels specialised for that model */
  const mpm_instance* inst;  /* device [n_inst] */
  const int32_t* slot_inst;  /* device [n_slots / 256]: instance owning each 256 slots */
  const int32_t* node_inst;  /* device [n_nodes / 256] */
  const int32_t* tile_inst;  /* device [n_tiles]: instance owning each unit */
  /* Optional (NULL: one launch of mixed-model kernels): device
   * [n_slots / 256], the 256-slot chunks grouped by model -- model m's chunks
   * are model_chunk[model_chunk_off[m] .. model_chunk_off[m + 1]).  The
   * model-dependent kernels then run once per model, specialised for it. */
  const int32_t* model_chunk;
  int64_t model_chunk_off[4];
} mpm_batch;

int32_t mpm_abi_version(void);
const char* mpm_last_error(void);

/* Bytes of scratch `ws` the step functions need for this batch and dtype
 * (0 = f32, 1 = f64).  Negative on error.  `ws` must be zero-filled once
 * before its first use; every sort leaves its bin counters and unit flags
 * cleared again, and a step that computes its sort keys from scratch clears
 * any a chained G2P left, so a workspace is reused across steps (one stream
 * at a time) without further clearing. */
int64_t mpm_workspace_bytes(const mpm_batch* b, int32_t dtype);

/* Tile shapes this build instantiates, as x,y,z triples; returns the count. */
int32_t mpm_tile_shapes(int32_t* out_xyz, int32_t max_shapes);

/* Constitutive seam over n particles (SoA [9][stride] per 3x3 field).
 * Replaces Model.stress (python/morphompm/constitutive.py:88-89,126-132). */
int32_t mpm_stress_f32(int32_t model, const double* prm6, int64_t n, int64_t stride,
                       const float* F, const float* Fg, const float* C, float* tau, void* stream);
int32_t mpm_stress_f64(int32_t model, const double* prm6, int64_t n, int64_t stride,
                       const double* F, const double* Fg, const double* C, double* tau,
                       void* stream);
/* Replaces Model.stress_vjp (python/morphompm/constitutive.py:91-93,134-147).
 * theta_bar ([2][stride], may be NULL) receives the Herschel-Bulkley
 * (tau_y, K) cotangents, an extension the reference does not return. */
int32_t mpm_stress_vjp_f32(int32_t model, const double* prm6, int64_t n, int64_t stride,
                           const float* F, const float* Fg, const float* C, const float* tau_bar,
                           float* F_bar, float* Fg_bar, float* C_bar, float* theta_bar,
                           void* stream);
int32_t mpm_stress_vjp_f64(int32_t model, const double* prm6, int64_t n, int64_t stride,
                           const double* F, const double* Fg, const double* C,
                           const double* tau_bar, double* F_bar, double* Fg_bar, double* C_bar,
                           double* theta_bar, void* stream);

/* Stable counting sort of particles by base cell (bin), fixed order inside a
 * bin (ascending slot).  Writes sortbuf.  No reference counterpart: the
 * reference loops particles in index order (python/morphompm/transfer.py:42). */
int32_t mpm_sort_by_cell_f32(const mpm_batch* b, const float* dyn, int32_t* sortbuf, void* ws,
                             void* stream);
int32_t mpm_sort_by_cell_f64(const mpm_batch* b, const double* dyn, int32_t* sortbuf, void* ws,
                             void* stream);

/* One forward step: sort, P2G (deterministic, no float atomics), grid update,
 * G2P.  Replaces transfer.step (python/morphompm/transfer.py:33-79); writes
 * dyn_out, and grid, sortbuf and payload, which step_bwd needs. */
int32_t mpm_step_fwd_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                         const float* sched, float* dyn_out, float* grid, int32_t* sortbuf,
                         float* payload, int32_t advect, void* ws, void* stream);
int32_t mpm_step_fwd_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                         const double* sched, double* dyn_out, double* grid, int32_t* sortbuf,
                         double* payload, int32_t advect, void* ws, void* stream);

/* mpm_step_fwd_f32 with no workspace argument, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * step's scratch is a 64 MiB array the library holds on each GPU it runs on
 * until the process ends; a batch whose forward workspace
 * (mpm_workspace_bytes up to the tile partials) exceeds it is rejected with
 * MPM_EINVAL before any device work.  Calls allocate nothing.  All calls on
 * one GPU share its array: calls that overlap on one GPU race.  Results are
 * bit for bit those of mpm_step_fwd_f32. */
int32_t mpm_step_fwd_noalloc_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                                 const float* sched, float* dyn_out, float* grid,
                                 int32_t* sortbuf, float* payload, int32_t advect,
                                 void* stream);

/* The same step for a chain of steps on one workspace and stream, with the
 * next step's P2G payload produced by this step's G2P (the per-particle prep
 * pass and the sort's position pass are then done once per chain, not once
 * per step).  ready = 1: `payload` and the sort keys in `ws` were written by
 * the previous chained call (its payload_next), which must be the last call
 * on this workspace; ready = 0: they are computed from dyn_in first.
 * payload_next (or NULL at the end of a chain) receives the next step's
 * payload, with the next step's growth factors sched_next.  out (MPM_OUT_*)
 * selects the rows of dyn_out written.  Results are bit for bit those of
 * mpm_step_fwd in the rows written.  No reference counterpart (a fusion of
 * consecutive transfer.step calls, python/morphompm/integrate.py:15-21). */
int32_t mpm_step_fwd_chain_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                               const float* sched, const float* sched_next, float* dyn_out,
                               float* grid, int32_t* sortbuf, float* payload, float* payload_next,
                               int32_t advect, int32_t ready, int32_t out, void* ws,
                               void* stream);
int32_t mpm_step_fwd_chain_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                               const double* sched, const double* sched_next, double* dyn_out,
                               double* grid, int32_t* sortbuf, double* payload,
                               double* payload_next, int32_t advect, int32_t ready, int32_t out,
                               void* ws, void* stream);

/* Reverse of one step.  Replaces transfer.step_vjp
 * (python/morphompm/transfer.py:82-158).  Inputs are the forward step's
 * dyn_in, Fg, sched, dyn_out, grid, sortbuf and payload, and the cotangents of the
 * step's outputs bar_out [24][n_slots].  Writes bar_in [24][n_slots] and adds
 * a * (local Fg cotangent) into Fg_bar [9][n_slots] (the reference's
 * pass-through Fg_b = Fgn_b + local when a = 1).  theta_acc (may be NULL)
 * accumulates the Herschel-Bulkley (tau_y, K) cotangents per 256-slot block,
 * [n_slots/256][2] float64. */
int32_t mpm_step_bwd_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                         const float* sched, const float* dyn_out, const float* grid,
                         const int32_t* sortbuf, const float* payload, const float* bar_out,
                         float* bar_in, float* Fg_bar, double* theta_acc, int32_t advect, void* ws,
                         void* stream);
int32_t mpm_step_bwd_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                         const double* sched, const double* dyn_out, const double* grid,
                         const int32_t* sortbuf, const double* payload, const double* bar_out,
                         double* bar_in, double* Fg_bar, double* theta_acc, int32_t advect,
                         void* ws, void* stream);

/* The same reverse step for a chain of reverse steps on one workspace and
 * stream (steps t, t-1, ... of one checkpoint segment).  ready = 0: this
 * step's G2P-adjoint payload is computed from dyn_in and bar_out first (the
 * first call of a chain); ready = 1 or 2: it is in ws from the previous
 * chained call (that of step t+1), which must be the last call on this
 * workspace -- the i-th call of a chain (i = 0, 1, ...) passes ready = 0 for
 * i = 0, else 1 + (i % 2): the workspace holds two payload areas and each
 * call reads one and writes the other.  dyn_prev / sortbuf_prev (both or
 * NULL): step t-1's input state and sort record; the call then also writes
 * step t-1's payload into ws, which carries C_bar, and the C rows of bar_in
 * are not written (the chain reads them from nowhere else; a caller that
 * needs C_bar of an intermediate step passes NULL).  Results are
 * bit for bit those of mpm_step_bwd in the rows written.  No reference
 * counterpart (a fusion of consecutive transfer.step_vjp calls,
 * python/morphompm/integrate.py:24-28). */
int32_t mpm_step_bwd_chain_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                               const float* sched, const float* dyn_out, const float* grid,
                               const int32_t* sortbuf, const float* payload, const float* bar_out,
                               float* bar_in, float* Fg_bar, double* theta_acc, int32_t advect,
                               int32_t ready, const float* dyn_prev, const int32_t* sortbuf_prev,
                               void* ws, void* stream);
int32_t mpm_step_bwd_chain_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                               const double* sched, const double* dyn_out, const double* grid,
                               const int32_t* sortbuf, const double* payload,
                               const double* bar_out, double* bar_in, double* Fg_bar,
                               double* theta_acc, int32_t advect, int32_t ready,
                               const double* dyn_prev, const int32_t* sortbuf_prev, void* ws,
                               void* stream);

/* ---- Batched growth inversion on the device -------------------------------
 * The caller side of a Gauss-Newton iteration over every instance of a batch,
 * so a whole iteration (growth, forward, loss, reverse, gradient, update) runs
 * without the host and can be captured as one graph.  Sums are in FP64 and in
 * a fixed order (per 256-slot chunk, then over an instance's chunks), so the
 * results are bitwise reproducible.  chunk_acc is caller-owned scratch of
 * n_slots / 256 doubles.  theta, loss, dloss, target: [n_inst] doubles.
 * dG: [9][n_slots] doubles (the growth direction of each slot). */

/* G = I + theta[i] dG for the live slots of instance i (I for padding).
 * The reference's parameterisation is Fg = g I
 * (python/morphompm/integrate.py:31-34 iso_growth_state); this is its
 * per-instance, per-particle-direction generalisation. */
int32_t mpm_growth_from_rates_f32(const mpm_batch* b, const double* theta, const double* dG,
                                  float* G, void* stream);
int32_t mpm_growth_from_rates_f64(const mpm_batch* b, const double* theta, const double* dG,
                                  double* G, void* stream);

/* loss[i] = sum over instance i's particles of det F (FP64), and the whole
 * cotangent state bar [24][n_slots]: det F F^-T in the F rows of live slots,
 * zero elsewhere.  Replaces loss_detF and the seed Ff_b of dloss_dg
 * (python/morphompm/integrate.py:41-42,49). */
int32_t mpm_loss_det_f32(const mpm_batch* b, const float* dyn, float* bar, double* chunk_acc,
                         double* loss, void* stream);
int32_t mpm_loss_det_f64(const mpm_batch* b, const double* dyn, double* bar, double* chunk_acc,
                         double* loss, void* stream);

/* grad[i] = sum over instance i's particles of <dG, G_bar> (FP64): dL/dtheta_i
 * for G = I + theta_i dG.  Replaces iso_growth_grad
 * (python/morphompm/integrate.py:37-38), which is the case dG = I. */
int32_t mpm_growth_grad_f32(const mpm_batch* b, const float* G_bar, const double* dG,
                            double* chunk_acc, double* grad, void* stream);
int32_t mpm_growth_grad_f64(const mpm_batch* b, const double* G_bar, const double* dG,
                            double* chunk_acc, double* grad, void* stream);

/* One damped Gauss-Newton iteration of every instance with done[i] == 0, the
 * loop body of infer_growth (python/morphompm/diff.py:14-23): |loss - target|
 * < tol sets done[i]; |dloss| < 1e-12 or non-finite nudges theta by 1e-3
 * towards the target; else theta -= clip((loss - target) / dloss, +-0.5).
 * iters[i] counts the iterations instance i took part in (the reference's
 * returned it + 1).  hist (optional, [n_inst][hist_cap]) receives each new
 * theta of a Gauss-Newton step at hist_len[i]++ while hist_len[i] < hist_cap
 * (the reference's history; the caller writes the initial value). */
int32_t mpm_gauss_newton(int32_t n_inst, const double* loss, const double* dloss,
                         const double* target, double tol, double* theta, int32_t* done,
                         int32_t* iters, double* hist, int32_t hist_cap, int32_t* hist_len,
                         void* stream);

/* Per-launch timing of one kernel (profiling aid, no reference counterpart).
 * `events` holds 2*n_pairs hipEvent_t handles created with timing enabled.
 * Every later launch of kernel `kernel_id` records events[2i] just before and
 * events[2i+1] just after it on its stream, for the i-th such launch while
 * i < n_pairs; launches under stream capture are not timed (event records
 * are refused inside a capture on this stack).  kernel_id < 0 disarms.
 * The timer is per host thread: it times the launches issued by the thread
 * that armed it, and calls from other threads neither see nor disturb it.
 * mpm_kernel_timer_count() = launches
 * recorded. */
enum {
  MPM_K_SORT = 0,        /* the whole counting-sort sequence */
  MPM_K_PREP_FWD = 1,
  MPM_K_P2G_FWD = 2,     /* forward P2G tile reduction */
  MPM_K_GRID_FWD = 3,
  MPM_K_G2P = 4,
  MPM_K_PREP_ADJ = 5,
  MPM_K_P2G_ADJ_TILE = 6, /* G2P-adjoint scatter as a tile reduction */
  MPM_K_GRID_ADJ = 7,
  MPM_K_ADJ_PARTICLE = 8 /* per particle: P2G-adjoint gather (x, v cotangents) and the
                            constitutive VJP (F, C, Fg cotangents) */
};
int32_t mpm_kernel_timer(int32_t kernel_id, void* const* events, int32_t n_pairs);
int32_t mpm_kernel_timer_count(void);

#ifdef __cplusplus
}
#endif

#endif /* MORPHOMPM_HIP_H */
