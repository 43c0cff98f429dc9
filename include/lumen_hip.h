/* lumen_hip.h -- the C ABI of the MI355X photon-transport engine (liblumen_hip.so).
 *
 * The engine replaces the reference's photon loop behind `lumen.simulate`: `run.trace(model, ops, rows, key, ...)`
 * plus `run.rebuild` (src/lumen/run.py:693-733, :534-557), which run tallymap's generated kernel over
 * `transport.step.step` (src/lumen/transport/step.py:516-571) and bin deposit records into per-sensor cells
 * (src/lumen/readout/fill.py:88-213).  Here one persistent HIP kernel traces every photon and accumulates the
 * per-sensor tallies directly in FP64, so no records leave the device.
 *
 * Conventions: every entry point returns LH_OK (0) or an error code, with a message from lh_last_error().  Pointers
 * named *_dev are device memory owned by the caller; everything else is host memory read during the call.  Calls on
 * one model are not thread-safe; a `stream` of NULL is the default stream.  Units: m, ns, nm, eV.
 */
#ifndef LUMEN_HIP_H
#define LUMEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LH_OK 0
#define LH_ERR_ARG 1    /* an argument is invalid (the message names it) */
#define LH_ERR_HIP 2    /* a HIP runtime call failed */
#define LH_ERR_LAYOUT 3 /* the packed layout is not this engine's */

/* source kinds: an isotropic point (src/lumen/sources/point.py:73-86), a beam, a Cherenkov segment
 * (src/lumen/sources/cherenkov.py:33-92) */
#define LH_SRC_POINT 0
#define LH_SRC_BEAM 1
#define LH_SRC_CHERENKOV 2

/* health cells (FP64), the reference's Health fates (src/lumen/api.py:445-495) plus counters */
#define LH_H_ABSORBED_BULK 0
#define LH_H_ABSORBED_SURFACE 1
#define LH_H_LOST 2
#define LH_H_ESCAPED 3
#define LH_H_DISTANCE_CUT 4
#define LH_H_ABANDONED 5
#define LH_H_TRUNCATED 6
#define LH_H_ROULETTED 7
#define LH_H_REWEIGHTED 8
#define LH_H_PUSHED 9
#define LH_H_RECORDS 10       /* deposits at a sensor (the reference's deposit records) */
#define LH_H_MOMENT_SPILLS 11 /* per-photon moment slots that overflowed (0 expected) */
#define LH_H_VERTICES 12      /* steps taken */
#define LH_H_PHOTONS 13       /* photons started */
#define LH_H_START_WEIGHT 14  /* their row weights summed */
#define LH_H_BOUNDED 15       /* steps whose query the reach bound stopped a few bands past the free path */
#define LH_HEALTH_N 16

typedef struct lh_model_s *lh_model;

/* The navigator arrays, by the reference's names (src/lumen/geometry/navigator_build.py:371-397 `host`,
 * src/lumen/geometry/csg_build.py:140-165 `T`, `K`, `G`).  Navigator ids: volumes 0..n_vol-1, the outside n_vol.
 * R, t: x_world = R x_local + t (R row-major 3x3).  A leaf is {(x-q)'A(x-q) + 2 b.(x-q) + c <= 0 (typ 1)}, or the
 * solid torus about the unit axis b through q with swept radius A[0] and tube radius A[1] (typ 2: the reference's
 * `R`, `r` columns), and {n_k.(x-o) <= h_k}; a kind's program is postfix (op 0 push leaf arg, 1 union, 2 inter,
 * 3 sub). */
typedef struct {
  int32_t n_vol, n_nodes, n_items, n_groups, n_leaves, n_planes, n_prog, n_kinds;
  const float *R;          /* n_vol*9 */
  const float *t;          /* n_vol*3 */
  const int32_t *kind;     /* n_vol */
  const int32_t *parent;   /* n_vol+1 (the outside's own entry is ignored) */
  const int32_t *cls;      /* n_vol+1 class per navigator id */
  const int32_t *sensor;   /* n_vol+1 sensor per navigator id, -1 none */
  const int32_t *vg0, *vg1; /* n_vol+1 each navigator id's frame groups [vg0, vg1) */
  const float *gR;         /* n_groups*9 each group's frame rotation */
  const int32_t *g0, *g1;  /* n_groups each group's BVH nodes [g0, g1) */
  const float *lo, *hi;    /* n_nodes*3 node boxes (group frame) */
  const int32_t *skip, *start, *count; /* n_nodes depth-first order; a leaf has count > 0 */
  const int32_t *items;    /* n_items daughter ids */
  const int32_t *leaf_typ, *leaf_p0, *leaf_np; /* n_leaves */
  const float *leaf_o, *leaf_q;                /* n_leaves*3 */
  const float *leaf_A;                         /* n_leaves*9 */
  const float *leaf_b;                         /* n_leaves*3 */
  const float *leaf_c;                         /* n_leaves */
  const float *plane_n;                        /* n_planes*3 */
  const float *plane_h;                        /* n_planes */
  const int32_t *kind_l0, *kind_nl, *kind_p0, *kind_pl; /* n_kinds */
  const int32_t *prog_op, *prog_arg;                    /* n_prog */
  float reach; /* the diagonal of the box holding every volume (m) */
  /* n_vol+1, or NULL: per volume the CSG leaf of its daughter-free core (-1 none), an acceleration with no
   * counterpart in the reference's arrays: no daughter meets that leaf's solid, so the walk skips every BVH box the
   * ray crosses inside it (lumen_amd/detector/navarrays.py skin_core) */
  const int32_t *skin;
  /* n_vol+1, or NULL (then every volume is taken to touch a sibling): 1 where a sibling's box lies within the
   * navigator's relocation step (K_ON float32 bands at the geometry's scale) of the volume's box, so an exit from the
   * volume may land in that sibling (touching siblings: the crossing goes on into it); no counterpart in the
   * reference's arrays (lumen_amd/detector/navarrays.py touch_flags) */
  const int32_t *touch;
  /* n_vol*3 each, or NULL (then no reach bound): each volume's world box (lo, hi), padded as the navigator's BVH
   * boxes (lumen_amd/detector/navarrays.py `lo_w`, `hi_w`).  lh_model_create builds the reach bound from them (the
   * reference's src/lumen/geometry/reach.py:174-222 `ReachGeometry`): the volumes' boxes padded by the relocation
   * step at the geometry's scale, and a BVH over the sensor volumes' boxes, so that a step whose free path ends
   * inside its volume's box before any sensor box along the ray queries no further than a few bands past it */
  const float *box_lo, *box_hi;
} lh_geometry;

/* One emission source.  Photon i of it (0 <= i < n) has the id first_id + i: its random numbers are a function of
 * (seed, id, vertex) only, so any split of a run into launches gives the same photons. */
typedef struct {
  int32_t kind;   /* LH_SRC_* */
  int32_t medium; /* Cherenkov: the medium whose RINDEX sets the cone and the spectrum's span */
  int32_t event;  /* the tally block this source adds into (0 <= event < n_events) */
  int32_t flags;
  float pos[3];   /* point position, or the segment's start (m) */
  float dir[3];   /* beam direction, or the segment's unit direction */
  float lam;      /* wavelength (nm) (point) */
  float lam_hi;   /* > lam: wavelengths uniform on [lam, lam_hi] (point) */
  float time;     /* emission time (ns) */
  float weight;   /* photons each Monte Carlo photon stands for (n_photons / samples), before Cherenkov's own */
  float length, beta, charge; /* Cherenkov segment */
  float pad;
  int64_t n;        /* photons of this source */
  int64_t first_id; /* the id of its first photon */
} lh_source;

/* Run controls (src/lumen/transport/step.py:52-71 Controls, src/lumen/run.py:70 MAX_VERTICES). */
typedef struct {
  float roulette;        /* epsilon > 0 */
  float max_distance;    /* <= 0: the geometry's reach */
  int32_t max_vertices;  /* the guard (steps per photon) */
  int32_t chunk;         /* photons a wavefront takes from the work queue at once (0: the engine's default) */
  int32_t blocks_per_cu; /* resident workgroups (768 threads) per CU of the persistent grid (0: as many as fit) */
  int32_t tally_global;  /* 1: accumulate the tallies with global atomics even where they fit in LDS; else 0 */
  int32_t nodes_global;  /* 1: read the BVH nodes from global memory even where they fit in LDS; else 0 */
  int32_t jump;          /* 1: the diffusion jump where the model holds a clearance grid (lh_model_jump); else 0 */
  int32_t keep_zero_weight; /* 1: trace a photon on after its live weight became exactly 0 (an edge row's floor
                             * drew an outcome the live values give no probability), as the reference does; 0: end
                             * it there.  It deposits nothing more either way: only the records and vertices
                             * diagnostics differ */
  int32_t reach_bound;      /* 1: the reach bound where the model has one (lh_geometry box_lo / box_hi); 0: every step
                             * queries to the next crossing.  The same photons and results either way; the bound
                             * pays where the query past a step's free path is costly (many volumes that are not
                             * sensors), and costs registers elsewhere */
} lh_controls;

/* Build a model: uploads the navigator arrays and the packed layout (lumen_amd/pack.py) to the current device.
 * Replaces the reference's Model + Operands build (src/lumen/transport/step.py:75-143). */
int lh_model_create(const lh_geometry *geometry, const int32_t *layout, int32_t layout_words, lh_model *out);
int lh_model_destroy(lh_model model);

/* The diffusion jump's clearance grid of medium `medium` (lumen_amd/detector/clearance.py): nodes lo + h i for
 * 0 <= i < n on each axis, `clear` (host, n[0] n[1] n[2] float32, index (i0 n[1] + i1) n[2] + i2) a proven lower
 * bound of the distance from each node to every surface.  A medium whose only channel is Rayleigh: at a Rayleigh
 * event in a launch whose controls set `jump`, the photon walks its own scatterings inside the ball of the clearance
 * at its position with no geometry query (the reference's src/lumen/transport/fastpath.py `diffuse`, with the grids
 * of its `prepare`, src/lumen/geometry/clearance.py).  Replaces an earlier grid of the medium; allocates. */
int lh_model_jump(lh_model model, int32_t medium, const float lo[3], float h, const int32_t n[3], const float *clear);

/* Build one BVH over n boxes (lo, hi: n*3 doubles, host): binned SAH with `bins` bins per axis, leaves of at most
 * `leaf_size` boxes, stored depth first with skip links (the navigator arrays' `lo`, `hi`, `skip`, `start`, `count`,
 * `items`; node_lo/node_hi hold 2n*3 doubles and skip/start/count 2n ints, items n).  Host only, no device call.
 * Replaces the reference's src/lumen/geometry/bvh.py:179-210 `build_many` (one range). */
int lh_bvh_build(const double *lo, const double *hi, int32_t n, int32_t leaf_size, int32_t bins, double *node_lo,
                 double *node_hi, int32_t *skip, int32_t *start, int32_t *count, int32_t *items, int32_t *n_nodes,
                 int32_t *depth);

/* Query the model: number of sensors, of charge cells per sensor (the QE wavelength knots times the CE theta knots,
 * the layout's words 5 and 14) and tally doubles per event. */
int lh_model_info(lh_model model, int32_t *n_sensors, int32_t *n_qe, int64_t *tally_per_event);

/* Trace every photon of `sources` (host array of n_sources) and add into `tally_dev` (n_events blocks of
 * [arrival S][arrival_t S][arrival2 S][charge S x n_qe (wavelength knot major, theta knot minor)], and with a time
 * grid in the layout (its words 11-13: nt,
 * t0, dt) [time S x nt][overflow S x (late q, late q t, early q, early q t)], FP64) and `health_dev` (LH_HEALTH_N
 * FP64).  Adds: zero
 * them first for a fresh run.  `values_dev`: the packed values (float32, device, n_values of them, each table
 * within them); `ref_values_dev`: the reference values every random choice is made at (NULL: values_dev); every
 * source's event is below n_events.  The call only enqueues work on `stream` (the sources go through the model's
 * pinned staging buffers): it returns before the kernel ends, and the model's calls are ordered by their stream,
 * so calls on one model use one stream (or the caller orders them).  Replaces run.trace + run.rebuild
 * (src/lumen/run.py:93-145). */
int lh_trace(lh_model model, const float *values_dev, const float *ref_values_dev, int64_t n_values,
             const lh_source *sources, int32_t n_sources, uint64_t seed, const lh_controls *controls,
             double *tally_dev, int32_t n_events, double *health_dev, void *stream);

/* The sources lh_trace_static stages at most (its scratch holds their copy and end offsets: 393 KiB a device). */
#define LH_STATIC_MAX_SOURCES 4096

/* lh_trace without allocating: the same kernels, arguments, results and errors, but the work-queue counter and the
 * staged sources live in a file-scope device array of fixed size (one copy per device, held until the process
 * ends; its address is taken once per device in lh_model_create), so the call makes no hipMalloc, hipFree or
 * pinned staging.  No counterpart in the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Limits: n_sources <= LH_STATIC_MAX_SOURCES
 * (else LH_ERR_ARG).  The sources are copied from host memory by a copy enqueued on `stream`: keep them unchanged
 * until the stream has passed the launch.  Every call on a device shares that device's array and leaves it as the
 * last call wrote it (the queue counter is cleared by the call itself): calls that overlap on one device race,
 * so order them on one stream. */
int lh_trace_static(lh_model model, const float *values_dev, const float *ref_values_dev, int64_t n_values,
                    const lh_source *sources, int32_t n_sources, uint64_t seed, const lh_controls *controls,
                    double *tally_dev, int32_t n_events, double *health_dev, void *stream);

/* Reverse-mode derivative by replay: each photon is traced twice from its Philox counter (no stored path): once to
 * total its cotangent-weighted deposits, once to add each step's score and pathwise terms times the deposits still
 * ahead.  `cotangent_dev`: dLoss/d(tally) in the tally layout (FP64); `slots_dev`: the gradient slot of each of the n_values packed
 * value (int32, -1 none, each below n_grad <= 8192); `grad_dev`: n_grad FP64 accumulators (added into); the
 * derivative is taken at `values_dev` with the draws at `ref_values_dev` (NULL: values_dev), the photons those of
 * lh_trace with the same seed and sources.  The arrival_t and arrival2 cells carry no cotangent (no derivative).  Replaces the reference's jvp/vjp
 * through tallymap's differentiated kernel (src/lumen/api.py:748-787). */
int lh_adjoint(lh_model model, const float *values_dev, const float *ref_values_dev, int64_t n_values,
               const int32_t *slots_dev, const lh_source *sources, int32_t n_sources, uint64_t seed,
               const lh_controls *controls, const double *cotangent_dev, int32_t n_events, double *grad_dev,
               int32_t n_grad, double *health_dev, void *stream);

/* A staged batch, for launches a caller captures in a HIP graph.  lh_batch_create uploads and checks the sources
 * (and, for the adjoint, the gradient slots: NULL slots_dev for a forward-only batch); it allocates and may
 * synchronise.  lh_batch_trace / lh_batch_adjoint then only enqueue kernels on `stream` (no allocation, no host
 * synchronisation), so a caller can capture them and replay the graph with new values and seeds written in place.
 * Each batch owns its sources and its device work queue: batches and lh_trace calls do not share state, but one
 * batch's launches must not overlap each other (they share its queue).  Replaces the reference's per-source
 * simulate loop over a Batch's rows (src/lumen/api.py:932-934, src/lumen/run.py:93-145) with one persistent
 * launch; the tallies are those of lh_trace over the same rows. */
typedef struct lh_batch_s *lh_batch;
int lh_batch_create(lh_model model, const lh_source *sources, int32_t n_sources, const int32_t *slots_dev,
                    int64_t n_values, int32_t n_grad, lh_batch *out, int64_t *total_photons);
int lh_batch_destroy(lh_batch batch);
int lh_batch_trace(lh_batch batch, const float *values_dev, const float *ref_values_dev, int64_t n_values,
                   uint64_t seed, const lh_controls *controls, double *tally_dev, int32_t n_events, double *health_dev,
                   void *stream);
int lh_batch_adjoint(lh_batch batch, const float *values_dev, const float *ref_values_dev, int64_t n_values,
                     uint64_t seed, const lh_controls *controls, const double *cotangent_dev, int32_t n_events,
                     double *grad_dev, int32_t n_grad, double *health_dev, void *stream);

/* The navigator on n rays (device arrays): from x (n*3) along d (n*3) in navigator id vol (-1: located here), the
 * photon sitting on face `last` (-1: none), the next visible crossing within tmax (<= 0: the reach): its distance
 * t, the navigator id entered (-1: none), the unit normal facing the ray (n*3), its surface code; `located` the
 * navigator id holding x.  Replaces Navigator.query / locate (src/lumen/geometry/navigate.py:663-693, :226-241)
 * as the reference's navigator tests call them (tests/unit/test_navigator.py:43-73). */
int lh_query(lh_model model, int32_t n, const float *x_dev, const float *d_dev, const int32_t *vol_dev,
             const int32_t *last_dev, float tmax, float *t_dev, int32_t *vol_after_dev, float *normal_dev,
             int32_t *srf_dev, int32_t *located_dev, void *stream);

/* The time of the last lh_trace / lh_adjoint's photon-loop kernel on its stream, in ms (HIP events); waits for that
 * kernel to finish. */
int lh_last_kernel_ms(lh_model model, float *ms);

/* Launch bookkeeping, so a caller times launches without waiting after each: lh_launches gives the number of
 * lh_trace / lh_adjoint launches made on the model so far (the next one is launch number *count), lh_kernel_ms_of
 * the kernel time of launch number `launch`, one of the last 256 (it waits for that kernel to finish).
 * Diagnostics: no counterpart in the reference (its runs are timed from the host). */
int lh_launches(lh_model model, int64_t *count);
int lh_kernel_ms_of(lh_model model, int64_t launch, float *ms);

/* The shape of the model's last launch (8 int32): grid blocks, resident blocks per CU, dynamic LDS bytes per
 * block, tallies in LDS (0/1), BVH nodes in LDS (0/1), the LDS budget per block, the CU's LDS bytes, the kernel's
 * feature bits (1 separate reference values, 2 a wavelength shifter, 4 a time histogram, 8 BVH nodes read from
 * memory, 16 torus leaves; 0x100 the adjoint).
 * Diagnostics: no counterpart in the reference. */
int lh_last_launch(lh_model model, int32_t *info);

const char *lh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LUMEN_HIP_H */
