/*
 * fastgeodis_hip.h -- C ABI of the MI355X (gfx950) raster-scan generalised
 * geodesic distance transform.
 *
 * Plain pointers and sizes only.  All pointers are device pointers on the
 * current HIP device; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Calls are asynchronous on `stream` and allocate nothing:
 * scratch memory comes from the caller's `workspace` (query its size with
 * fg_workspace_bytes).  Calls may come from several host threads on different
 * streams, each with its own workspace.  The tiled sweep kernels need all of
 * a launch's workgroups resident at once, so the library orders the work of
 * such calls per device: a call on a different stream than the previous one
 * waits (on the device) for that call's end.  The only shared state is that
 * order and read-only caches.
 *
 * Layouts (batch 1, C-contiguous float32):
 *   image (C, [D,] H, W), mask (Cm, [D,] H, W), out (Cm, [D,] H, W).
 * Output channel 0 holds the transform; channels >= 1 hold the same
 * elementwise function of the mask that the reference returns for them
 * (reference fastgeodis_cpu.cpp:174, 356 initialise every channel, the passes
 * touch channel 0 only).
 *
 * Return codes map onto the reference's Python exceptions:
 *   FG_OK            0
 *   FG_EINVAL       -1  invalid argument        -> ValueError
 *   FG_EHIP         -2  HIP runtime error       -> RuntimeError
 *   FG_EWATCHDOG    -3  a bounded device wait expired (a tile never arrived)
 *                       -> RuntimeError
 *   FG_EWORKSPACE   -4  workspace too small     -> RuntimeError
 * fg_last_error() returns a thread-local message for the last failure.
 */
#ifndef FASTGEODIS_HIP_H
#define FASTGEODIS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_OK 0
#define FG_EINVAL (-1)
#define FG_EHIP (-2)
#define FG_EWATCHDOG (-3)
#define FG_EWORKSPACE (-4)

/* Operation kinds for fg_workspace_bytes and fg_instance.op. */
#define FG_OP_UNSIGNED 0
#define FG_OP_SIGNED 1
#define FG_OP_GSF 2

/* ABI version of this header (bumped on any incompatible change). */
int fg_abi_version(void);

/* Thread-local text of the last error returned on this host thread. */
const char *fg_last_error(void);

/* Bytes of device workspace one call of kind `op` needs for an image of C
 * channels over D*H*W voxels; D <= 0 selects the 2D entry points (H*W).  Depends on the tuning
 * (fg_tuning_set) of the calling host thread. */
size_t fg_workspace_bytes(int op, int C, int D, int H, int W);

/* ---- device status (no counterpart in the reference, whose kernels report
 * nothing, fastgeodis_cuda.cu:232-260, 541-570).  Each device has one
 * persistent status word in pinned host memory.  A launch sets it when a
 * bounded device-side wait expires (FG_EWATCHDOG: a tile never arrived, the
 * results of that call are invalid) or a launch is refused.  Every compute
 * call first reads and clears the word of its device, without
 * synchronising, and fails with that code when it was set, so an expiry is
 * reported by the next call at the latest.
 *
 * fg_status synchronises `stream` and then reads and clears the word of the
 * current device: FG_OK, FG_EWATCHDOG or FG_EHIP.  fg_pending_status reads
 * and clears the word of `device` without any HIP call.  fg_inject_status
 * sets a device's pending status from the host (tests of the error path).
 * fg_set_wait_limit sets how many checks of a pending halo value a bounded
 * wait makes before it expires, for later calls on this host thread (0 = the
 * default, about 2^21; 1 expires at the first check that finds a value
 * missing). */
int fg_status(void *stream);
int fg_pending_status(int device);
int fg_inject_status(int device, int code);
int fg_set_wait_limit(int spins);

/* Waits for outstanding work of the library's events and frees its pinned
 * host memory and events (call once at process exit; later calls
 * re-create what they need). */
void fg_release(void);

/* Replaces generalised_geodesic2d_cuda (reference fastgeodis.h:38-44,
 * fastgeodis_cuda.cu:278-318): out = transform of fl(v*mask) after
 * `iterations` iterations of (row pass pair, column pass pair).
 * l_grad = lambda and l_eucl = 1 - lambda as the reference's Python wrapper
 * passes them (FastGeodis/__init__.py:63-65). */
int fg_geodesic2d_f32(const float *image, const float *mask, float *out, int C,
                      int Cm, int H, int W, float v, float l_grad,
                      float l_eucl, int iterations, void *workspace,
                      size_t workspace_bytes, void *stream);

/* Replaces generalised_geodesic3d_cuda (reference fastgeodis.h:46-54,
 * fastgeodis_cuda.cu:587-656).  spacing = (s_D, s_H, s_W), unsquared; it is
 * squared in float and the per-pass step tables are built in the reference's
 * per-pass summation order (fastgeodis_cpu.cpp:220-231, 353, 365-397). */
int fg_geodesic3d_f32(const float *image, const float *mask, float *out, int C,
                      int Cm, int D, int H, int W, const float spacing[3],
                      float v, float l_grad, float l_eucl, int iterations,
                      void *workspace, size_t workspace_bytes, void *stream);

/* Replaces signed_generalised_geodesic2d/3d (reference fastgeodis.cpp:241-258):
 * out = D(mask) - D(1 - mask).  Both transforms run in the same launches. */
int fg_signed_geodesic2d_f32(const float *image, const float *mask, float *out,
                             int C, int Cm, int H, int W, float v,
                             float l_grad, float l_eucl, int iterations,
                             void *workspace, size_t workspace_bytes,
                             void *stream);
int fg_signed_geodesic3d_f32(const float *image, const float *mask, float *out,
                             int C, int Cm, int D, int H, int W,
                             const float spacing[3], float v, float l_grad,
                             float l_eucl, int iterations, void *workspace,
                             size_t workspace_bytes, void *stream);

/* Replaces GSF2d/GSF3d (reference fastgeodis.cpp:319-343).  `lambda` is the
 * raw lamb; l_eucl = 1 - lambda is formed in float as the reference does. */
int fg_gsf2d_f32(const float *image, const float *mask, float *out, int C,
                 int Cm, int H, int W, float theta, float v, float lambda,
                 int iterations, void *workspace, size_t workspace_bytes,
                 void *stream);
int fg_gsf3d_f32(const float *image, const float *mask, float *out, int C,
                 int Cm, int D, int H, int W, float theta,
                 const float spacing[3], float v, float lambda, int iterations,
                 void *workspace, size_t workspace_bytes, void *stream);

/* ---- no-allocation variant of fg_geodesic3d_f32.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as fg_geodesic3d_f32 without the workspace: the
 * scratch (halo exchange, transposed image and distance) is a static device
 * array of fg_noalloc_scratch_bytes() bytes (256 MiB less 256 B, about 33 M
 * voxels at C = 1) that each device holds from the library's first use on it
 * until the process ends; larger volumes return FG_EWORKSPACE.  It calls no
 * hipMalloc, hipFree, host-pinned allocation or event creation.  All calls of
 * this entry point on one device share that array: calls that overlap on one
 * device (other streams, other host threads) race and must be serialised by
 * the caller.  Device errors are reported like the other entry points'; while
 * no other entry point has run on the device, only fg_status reports them. */
int fg_geodesic3d_f32_noalloc(const float *image, const float *mask,
                              float *out, int C, int Cm, int D, int H, int W,
                              const float spacing[3], float v, float l_grad,
                              float l_eucl, int iterations, void *stream);
size_t fg_noalloc_scratch_bytes(void);

/* ---- batched form (no counterpart in the reference, which accepts batch 1
 * only, common.h:89-95): many instances of uneven sizes, 2D (D = 1) and 3D
 * mixed, in the same persistent launches.  Each instance's result equals the
 * single-instance call with the same arguments bit for bit. */
typedef struct fg_instance {
  const float *image; /* (C, D, H, W) */
  const float *mask;  /* (Cm, D, H, W) */
  float *out;         /* (Cm, D, H, W) */
  int C, Cm, D, H, W;
  int is2d;           /* 1: 2D semantics (D must be 1, spacing ignored) */
  int op;             /* FG_OP_UNSIGNED / FG_OP_SIGNED / FG_OP_GSF */
  int iterations;
  float spacing[3];   /* unsquared (s_D, s_H, s_W); 3D only */
  float v;
  float l_grad;       /* unsigned/signed: lambda; GSF: lambda */
  float l_eucl;       /* unsigned/signed: 1 - lambda; GSF: ignored */
  float theta;        /* GSF only */
  const float *imageT; /* optional (else NULL): the image already transposed
                          to (C, W, H, D) (3D) or (C, W, H) (2D), e.g. kept by
                          a caller that transforms one image many times; the
                          call then skips its image transpose and workspace */
} fg_instance;

size_t fg_batch_workspace_bytes(const fg_instance *inst, int n);
int fg_geodesic_batch_f32(const fg_instance *inst, int n, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Stepper: continuous batching at pass granularity (no counterpart in the
 * reference; it serves the click-serving loop).  Unsigned and signed
 * transforms (iterations >= 1, first sweep on the tiled engine) are added at
 * any time; each fg_stepper_step records, for every active transform, its
 * next pass pair (forward and backward sweep) in launches shared with every
 * other active transform at that pass, plus the transposes around W passes
 * that its sweeps do not replace, and enqueues them on `stream`.  A transform whose last W pass ran in a
 * step is reported in `done` (its ticket) and has left the stepper; its
 * output is complete when `stream` reaches the end of that step.  Results
 * equal the same instance run alone.
 *   halo: device memory for the halo exchange of all active transforms
 *         (fg_stepper_op_bytes reports each transform's share; fg_stepper_add
 *         refuses a transform that does not fit, FG_EWORKSPACE);
 *   blob: device memory for one step's descriptors, fg_stepper_blob_bytes(
 *         maximum active transforms; a signed instance counts 2);
 *   workspace (per instance, fg_stepper_op_bytes): kept by the caller until
 *         the instance is reported done, like its image, mask and output.
 * The mask is read by the instance's first step only.  A stepper is used by
 * one host thread at a time, on the device current at its creation. */
typedef struct fg_stepper fg_stepper;
size_t fg_stepper_blob_bytes(int max_transforms);
size_t fg_stepper_op_bytes(const fg_instance *inst, size_t *halo_bytes);
int fg_stepper_create(fg_stepper **out, void *halo, size_t halo_bytes, void *blob,
                      size_t blob_bytes);
int fg_stepper_add(fg_stepper *s, const fg_instance *inst, void *workspace,
                   size_t workspace_bytes, int *ticket);
int fg_stepper_step(fg_stepper *s, void *stream, int *done, int cap, int *ndone);
int fg_stepper_active(const fg_stepper *s);
void fg_stepper_destroy(fg_stepper *s);

/* Writes n single float values (items[i].dst[items[i].index] = value),
 * asynchronously on `stream`, in one launch per 96 items; the items are
 * passed by value (no device allocation).  No counterpart in the reference:
 * it applies click edits to resident masks in the serving loop. */
typedef struct fg_voxel_set {
  float *dst;
  long long index;
  float value;
} fg_voxel_set;
int fg_set_voxels_f32(const fg_voxel_set *items, int n, void *stream);

/* Tile/launch parameters of the persistent sweep engine.  engine 0 is the
 * one-launch-per-slice engine, 1 the persistent tiled engine (default).
 * halo is the trapezoid depth K (slices per neighbour exchange, 1..30),
 * rounded up to a multiple of the kernel's slab ring; 0 (default) picks the
 * deepest halo whose tiles fit one per compute unit: per 2D row and per lone
 * 3D plane, and over all 3D planes that share a pass (batches, the halves of
 * a signed or GSF call).  Values are per host thread; fg_tuning_get reads
 * them back. */
typedef struct fg_tuning {
  int engine;
  int halo;
  int max_groups; /* cap on persistent workgroups (0 = occupancy limit) */
  int shape;      /* 3D plane workgroup shape (waves x rows / slab ring): 0 auto
                     (= 5), 1 = 4x8/2, 2 = 8x4/3, 3 = 16x2/3, 4 = 8x4/5, 5 = 16x2/5,
                     6 = 8x4/5 plus four waves that perform the halo exchanges
                     (1-2 channels; 3-4 channels run 8x4/2), 7 = 8x2/5 (16-row
                     windows; 3-4 channels 8x2/3) */
} fg_tuning;
int fg_tuning_set(const fg_tuning *t);
int fg_tuning_get(fg_tuning *t);

/* Graph replay (no counterpart in the reference, whose host loop launches
 * every pass, fastgeodis_cuda.cu:290-313, 604-646).  While on (the default,
 * per host thread), a call whose sequence of launches repeats one seen before
 * on this thread and device runs as a single hipGraphLaunch of a cached
 * executable graph (kernel nodes only) updated with this call's arguments;
 * results are identical either way.  Off: every launch is issued directly. */
int fg_set_graphs(int on);

/* Per-launch timing of the persistent sweep kernels, for measurement tools
 * (no counterpart in the reference).  While enabled on a host thread, every
 * tiled sweep launch it issues is bracketed by HIP events on its stream
 * (fg_profile_enable(1); 2 records the exchange-wait histogram below instead,
 * 3 both; 0 off).
 * fg_profile_collect waits for those events and returns the number of
 * launches, their summed duration (ms) and their summed algorithmic bytes
 * ((4*C + 8) bytes per voxel update), then clears the record. */
int fg_profile_enable(int on);
int fg_profile_collect(int *launches, double *total_ms, double *alg_bytes);

/* While the histogram is enabled (fg_profile_enable(2)), every wave of every
 * tiled sweep records
 * how long each of its halo exchanges took (from its publications to the
 * last neighbour value it waited for) on the device's constant-rate clock
 * (hipDeviceAttributeWallClockRate kHz), as a histogram of 32 log2 buckets:
 * bucket b counts exchanges of [2^(b-1), 2^b) ticks, bucket 0 of 0 ticks.
 * Each workgroup counts in its LDS and adds its counts to the device's
 * histogram once, at its end.  fg_profile_enable(2) clears the current
 * device's histogram;
 * fg_profile_exchange_histogram synchronises the device and copies it into
 * counts[0..31] (n >= 32). */
int fg_profile_exchange_histogram(unsigned *counts, int n);

#ifdef __cplusplus
}
#endif

#endif /* FASTGEODIS_HIP_H */
