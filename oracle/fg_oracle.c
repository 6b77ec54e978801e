/*
 * fg_oracle.c -- CPU restatement of FastGeodis' raster-scan generalised
 * geodesic distance transform (reference v1.0.5).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path (fastgeodis_amd) never calls it.
 *
 * Semantics followed (reference file:line):
 *   - init  d = fl(v * mask) for every mask channel      fastgeodis_cpu.cpp:174, 356
 *   - 2D pass pair (top-down then bottom-up along rows)  fastgeodis_cpu.cpp:53-162
 *   - 2D driver: rows, then columns (via transpose)      fastgeodis_cpu.cpp:164-201
 *   - 3D pass pair with per-pass 3x3 step table          fastgeodis_cpu.cpp:203-340
 *   - 3D driver: spacing squared in float, D/H/W passes  fastgeodis_cpu.cpp:342-407
 *   - l1distance |a-b|, channel sum from 0.0f            fastgeodis_cpu.cpp:38-41, 88-105
 *   - signed = D(M) - D(1-M)                             fastgeodis.cpp:241-258
 *   - GSF: three signed transforms, thresholds +-theta   fastgeodis.cpp:319-343
 *   - Toivanen's raster scan (accuracy cross-check)      geodis_toivanen.cpp:40-348
 *
 * The reference transposes tensors so that every sweep runs along the
 * outermost spatial axis; transposes only move data, so this restatement
 * indexes the native layout with explicit strides instead.  Each candidate is
 * evaluated as (d_q + l_eucl*LD) + l_grad*l with two float roundings and no
 * fused multiply-add (build with -ffp-contract=off), and the running minimum
 * keeps the current value unless the candidate is strictly smaller, which is
 * std::min's rule.  Neighbours are visited in the reference order (first
 * in-plane axis outer, second inner).
 *
 * The per-slice loops are OpenMP-parallel like the reference (one parallel
 * region per row or plane), so the same code serves as a CPU baseline.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min(a, b): returns b only when b < a (fastgeodis_cpu.cpp:110). */
static inline float ref_min(float a, float b) { return (b < a) ? b : a; }

/*
 * One pass pair (forward sweep then backward sweep) along axis A of one
 * distance channel.  Slices along A are updated in order; every element of
 * slice a is recomputed from the already-updated slice a-1 (forward) or a+1
 * (backward) over the 3x3 neighbourhood spanned by in-plane axes (B, E).
 * Out-of-range neighbours are skipped.  ld[i*3+j] is the step length for
 * offset (i-1, j-1) in (B, E).  Mirrors fastgeodis_cpu.cpp:234-339 (3D) and,
 * with nB == 1 and ld row 1 = {sqrt2, 1, sqrt2}, fastgeodis_cpu.cpp:70-161 (2D).
 */
static void pass_pair(const float *img, float *d, int C, ptrdiff_t cs,
                      int nA, ptrdiff_t sA, int nB, ptrdiff_t sB, int nE,
                      ptrdiff_t sE, const float ld[9], float l_grad,
                      float l_eucl) {
  for (int dir = 0; dir < 2; ++dir) {
    const int step = dir == 0 ? 1 : -1;
    const int first = dir == 0 ? 1 : nA - 2;
    for (int a = first; a >= 0 && a < nA; a += step) {
      const int ap = a - step;
#pragma omp parallel for collapse(2) schedule(static)
      for (int b = 0; b < nB; ++b) {
        for (int e = 0; e < nE; ++e) {
          const ptrdiff_t p = (ptrdiff_t)a * sA + (ptrdiff_t)b * sB + (ptrdiff_t)e * sE;
          float new_dist = d[p];
          for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
              const int bq = b + i - 1, eq = e + j - 1;
              if (bq < 0 || bq >= nB || eq < 0 || eq >= nE) continue;
              const ptrdiff_t q = (ptrdiff_t)ap * sA + (ptrdiff_t)bq * sB + (ptrdiff_t)eq * sE;
              float l_dist = 0.0f;
              if (C == 1) {
                l_dist = fabsf(img[p] - img[q]);
              } else {
                for (int c = 0; c < C; ++c)
                  l_dist += fabsf(img[(ptrdiff_t)c * cs + p] - img[(ptrdiff_t)c * cs + q]);
              }
              const float cur = d[q] + l_eucl * ld[i * 3 + j] + l_grad * l_dist;
              new_dist = ref_min(new_dist, cur);
            }
          }
          d[p] = new_dist;
        }
      }
    }
  }
}

/* Step table of one 3D pass: sqrt((s0 + |i-1|*s1) + |j-1|*s2), summed in this
 * order in float (fastgeodis_cpu.cpp:220-231).  s* are squared spacings. */
static void table3d(float s0, float s1, float s2, float ld[9]) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      float v = s0;
      v += (float)abs(i - 1) * s1;
      v += (float)abs(j - 1) * s2;
      ld[i * 3 + j] = sqrtf(v);
    }
  }
}

static void init_dist(const float *mask, float *out, size_t n, float v) {
  for (size_t i = 0; i < n; ++i) out[i] = v * mask[i];
}

/* Runs the 2D iterations on distance channel 0 (fastgeodis_cpu.cpp:177-198). */
static void transform2d(const float *img, float *d, int C, int H, int W,
                        float l_grad, float l_eucl, int iters) {
  const float r2 = sqrtf(2.0f);
  const float ld[9] = {0.f, 0.f, 0.f, r2, 1.0f, r2, 0.f, 0.f, 0.f};
  const ptrdiff_t cs = (ptrdiff_t)H * W;
  for (int it = 0; it < iters; ++it) {
    /* rows: sweep along H, neighbours along W */
    pass_pair(img, d, C, cs, H, W, 1, 0, W, 1, ld, l_grad, l_eucl);
    /* columns: sweep along W, neighbours along H */
    pass_pair(img, d, C, cs, W, 1, 1, 0, H, W, ld, l_grad, l_eucl);
  }
}

/* Runs the 3D iterations on distance channel 0 (fastgeodis_cpu.cpp:359-404).
 * sq[] holds the squared spacings (D, H, W). */
static void transform3d(const float *img, float *d, int C, int D, int H, int W,
                        const float sq[3], float l_grad, float l_eucl,
                        int iters) {
  float ldD[9], ldH[9], ldW[9];
  table3d(sq[0], sq[1], sq[2], ldD); /* pass along D: in-plane (H, W) */
  table3d(sq[1], sq[0], sq[2], ldH); /* pass along H: in-plane (D, W) */
  table3d(sq[2], sq[1], sq[0], ldW); /* pass along W: in-plane (H, D) */
  const ptrdiff_t sD = (ptrdiff_t)H * W, sH = W, sW = 1;
  const ptrdiff_t cs = (ptrdiff_t)D * H * W;
  for (int it = 0; it < iters; ++it) {
    pass_pair(img, d, C, cs, D, sD, H, sH, W, sW, ldD, l_grad, l_eucl);
    pass_pair(img, d, C, cs, H, sH, D, sD, W, sW, ldH, l_grad, l_eucl);
    pass_pair(img, d, C, cs, W, sW, H, sH, D, sD, ldW, l_grad, l_eucl);
  }
}

/* ---- public entry points (ctypes) ---------------------------------------
 * Shapes: image (1, C, [D,] H, W), mask/out (1, Cm, [D,] H, W), contiguous.
 * Output channel 0 is transformed; channels >= 1 hold fl(v * mask). */

int fgo_geodesic2d(const float *img, const float *mask, float *out, int C,
                   int Cm, int H, int W, float v, float l_grad, float l_eucl,
                   int iters) {
  if (C < 1 || Cm < 1 || H < 1 || W < 1) return -1;
  init_dist(mask, out, (size_t)Cm * H * W, v);
  transform2d(img, out, C, H, W, l_grad, l_eucl, iters);
  return 0;
}

int fgo_geodesic3d(const float *img, const float *mask, float *out, int C,
                   int Cm, int D, int H, int W, const float spacing[3], float v,
                   float l_grad, float l_eucl, int iters) {
  if (C < 1 || Cm < 1 || D < 1 || H < 1 || W < 1) return -1;
  float sq[3];
  for (int k = 0; k < 3; ++k) sq[k] = spacing[k] * spacing[k];
  init_dist(mask, out, (size_t)Cm * D * H * W, v);
  transform3d(img, out, C, D, H, W, sq, l_grad, l_eucl, iters);
  return 0;
}

/* spatial dims packed as dims[0..nd-1]; D = 1 for 2D. */
static int unsigned_nd(const float *img, const float *mask, float *out, int C,
                       int Cm, int nd, const int *dims, const float *spacing,
                       float v, float l_grad, float l_eucl, int iters) {
  if (nd == 2)
    return fgo_geodesic2d(img, mask, out, C, Cm, dims[0], dims[1], v, l_grad,
                          l_eucl, iters);
  return fgo_geodesic3d(img, mask, out, C, Cm, dims[0], dims[1], dims[2],
                        spacing, v, l_grad, l_eucl, iters);
}

/* signed = D(M) - D(1 - M)  (fastgeodis.cpp:241-258) */
static int signed_nd(const float *img, const float *mask, float *out, int C,
                     int Cm, int nd, const int *dims, const float *spacing,
                     float v, float l_grad, float l_eucl, int iters) {
  size_t n = (size_t)Cm;
  for (int k = 0; k < nd; ++k) n *= (size_t)dims[k];
  float *inv = (float *)malloc(n * sizeof(float));
  float *db = (float *)malloc(n * sizeof(float));
  if (!inv || !db) {
    free(inv);
    free(db);
    return -2;
  }
  for (size_t i = 0; i < n; ++i) inv[i] = 1.0f - mask[i];
  int rc = unsigned_nd(img, mask, out, C, Cm, nd, dims, spacing, v, l_grad,
                       l_eucl, iters);
  if (rc == 0)
    rc = unsigned_nd(img, inv, db, C, Cm, nd, dims, spacing, v, l_grad, l_eucl,
                     iters);
  if (rc == 0)
    for (size_t i = 0; i < n; ++i) out[i] = out[i] - db[i];
  free(inv);
  free(db);
  return rc;
}

/* GSF (fastgeodis.cpp:319-343): l_eucl = 1 - lambda in float. */
static int gsf_nd(const float *img, const float *mask, float *out, int C,
                  int Cm, int nd, const int *dims, const float *spacing,
                  float theta, float v, float lambda, int iters) {
  size_t n = (size_t)Cm;
  for (int k = 0; k < nd; ++k) n *= (size_t)dims[k];
  const float l_eucl = 1 - lambda;
  float *ds = (float *)malloc(n * sizeof(float));
  float *m = (float *)malloc(n * sizeof(float));
  float *dd = (float *)malloc(n * sizeof(float));
  if (!ds || !m || !dd) {
    free(ds);
    free(m);
    free(dd);
    return -2;
  }
  int rc = signed_nd(img, mask, ds, C, Cm, nd, dims, spacing, v, lambda,
                     l_eucl, iters);
  if (rc == 0) {
    /* Dd_Md = -signed(1 - Md), Md = (Ds > theta) */
    for (size_t i = 0; i < n; ++i) m[i] = 1.0f - ((ds[i] > theta) ? 1.0f : 0.0f);
    rc = signed_nd(img, m, dd, C, Cm, nd, dims, spacing, v, lambda, l_eucl,
                   iters);
  }
  if (rc == 0) {
    /* De_Me = signed(Me), Me = (Ds > -theta) */
    for (size_t i = 0; i < n; ++i) m[i] = (ds[i] > -theta) ? 1.0f : 0.0f;
    rc = signed_nd(img, m, out, C, Cm, nd, dims, spacing, v, lambda, l_eucl,
                   iters);
  }
  if (rc == 0)
    for (size_t i = 0; i < n; ++i) out[i] = -dd[i] + out[i];
  free(ds);
  free(m);
  free(dd);
  return rc;
}

int fgo_signed_geodesic2d(const float *img, const float *mask, float *out,
                          int C, int Cm, int H, int W, float v, float l_grad,
                          float l_eucl, int iters) {
  const int dims[2] = {H, W};
  return signed_nd(img, mask, out, C, Cm, 2, dims, NULL, v, l_grad, l_eucl,
                   iters);
}

int fgo_signed_geodesic3d(const float *img, const float *mask, float *out,
                          int C, int Cm, int D, int H, int W,
                          const float spacing[3], float v, float l_grad,
                          float l_eucl, int iters) {
  const int dims[3] = {D, H, W};
  return signed_nd(img, mask, out, C, Cm, 3, dims, spacing, v, l_grad, l_eucl,
                   iters);
}

int fgo_gsf2d(const float *img, const float *mask, float *out, int C, int Cm,
              int H, int W, float theta, float v, float lambda, int iters) {
  const int dims[2] = {H, W};
  return gsf_nd(img, mask, out, C, Cm, 2, dims, NULL, theta, v, lambda, iters);
}

int fgo_gsf3d(const float *img, const float *mask, float *out, int C, int Cm,
              int D, int H, int W, float theta, const float spacing[3],
              float v, float lambda, int iters) {
  const int dims[3] = {D, H, W};
  return gsf_nd(img, mask, out, C, Cm, 3, dims, spacing, theta, v, lambda,
                iters);
}

int fgo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Sets the OpenMP thread count of later calls (n >= 1). */
void fgo_set_num_threads(int n) {
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- layout-faithful forms (the CPU baseline) ----------------------------
 * The reference runs every pass on a contiguous copy whose outermost spatial
 * axis is the sweep axis: per 3D iteration it materialises the image and the
 * distance in (H, D, W) for the H pass, in (W, H, D) for the W pass and back
 * in (D, H, W) (transpose + contiguous at fastgeodis_cpu.cpp:361-401); per 2D
 * iteration in (W, H) for the column pass and back (:180-195).  These forms
 * do the same copies, OpenMP-parallel, and run the same pass_pair on
 * contiguous rows, so their timing reflects the reference's memory traffic.
 * Results equal fgo_geodesic2d/3d bit for bit (copies only move data). */

/* dst (C, n0, n1, n2) contiguous; dst[c][i][j][k] = src[c*cs + i*s0 + j*s1 + k*s2] */
static void permute_copy(const float *src, float *dst, int C, ptrdiff_t cs,
                         int n0, int n1, int n2, ptrdiff_t s0, ptrdiff_t s1,
                         ptrdiff_t s2) {
  const int T = 32;
#pragma omp parallel for collapse(2) schedule(static)
  for (int c = 0; c < C; ++c) {
    for (int i = 0; i < n0; ++i) {
      const float *sp = src + (ptrdiff_t)c * cs + (ptrdiff_t)i * s0;
      float *dp = dst + ((ptrdiff_t)c * n0 + i) * n1 * (ptrdiff_t)n2;
      for (int j0 = 0; j0 < n1; j0 += T)
        for (int k0 = 0; k0 < n2; k0 += T)
          for (int j = j0; j < j0 + T && j < n1; ++j)
            for (int k = k0; k < k0 + T && k < n2; ++k)
              dp[(ptrdiff_t)j * n2 + k] = sp[(ptrdiff_t)j * s1 + (ptrdiff_t)k * s2];
    }
  }
}

int fgo_geodesic3d_layout(const float *img, const float *mask, float *out,
                          int C, int Cm, int D, int H, int W,
                          const float spacing[3], float v, float l_grad,
                          float l_eucl, int iters) {
  if (C < 1 || Cm < 1 || D < 1 || H < 1 || W < 1) return -1;
  const size_t n = (size_t)D * H * W;
  float sq[3];
  for (int k = 0; k < 3; ++k) sq[k] = spacing[k] * spacing[k];
  float ldD[9], ldH[9], ldW[9];
  table3d(sq[0], sq[1], sq[2], ldD);
  table3d(sq[1], sq[0], sq[2], ldH);
  table3d(sq[2], sq[1], sq[0], ldW);
  init_dist(mask, out, (size_t)Cm * n, v);
  float *ia = (float *)malloc((size_t)C * n * sizeof(float));
  float *ib = (float *)malloc((size_t)C * n * sizeof(float));
  float *da = (float *)malloc(n * sizeof(float));
  float *db = (float *)malloc(n * sizeof(float));
  if (!ia || !ib || !da || !db) {
    free(ia); free(ib); free(da); free(db);
    return -2;
  }
  const ptrdiff_t cs = (ptrdiff_t)n;
  /* image_local = image.clone(); distance channel 0 in da (D, H, W) */
  memcpy(ia, img, (size_t)C * n * sizeof(float));
  memcpy(da, out, n * sizeof(float));
  for (int it = 0; it < iters; ++it) {
    /* D pass on (D, H, W) */
    pass_pair(ia, da, C, cs, D, (ptrdiff_t)H * W, H, W, W, 1, ldD, l_grad, l_eucl);
    /* (D, H, W) -> (H, D, W) */
    permute_copy(ia, ib, C, cs, H, D, W, W, (ptrdiff_t)H * W, 1);
    permute_copy(da, db, 1, cs, H, D, W, W, (ptrdiff_t)H * W, 1);
    pass_pair(ib, db, C, cs, H, (ptrdiff_t)D * W, D, W, W, 1, ldH, l_grad, l_eucl);
    /* (H, D, W) -> (W, H, D) */
    permute_copy(ib, ia, C, cs, W, H, D, 1, (ptrdiff_t)D * W, W);
    permute_copy(db, da, 1, cs, W, H, D, 1, (ptrdiff_t)D * W, W);
    pass_pair(ia, da, C, cs, W, (ptrdiff_t)H * D, H, D, D, 1, ldW, l_grad, l_eucl);
    /* (W, H, D) -> (D, H, W) */
    permute_copy(ia, ib, C, cs, D, H, W, 1, D, (ptrdiff_t)H * D);
    permute_copy(da, db, 1, cs, D, H, W, 1, D, (ptrdiff_t)H * D);
    float *t = ia; ia = ib; ib = t;
    t = da; da = db; db = t;
  }
  memcpy(out, da, n * sizeof(float));
  free(ia); free(ib); free(da); free(db);
  return 0;
}

int fgo_geodesic2d_layout(const float *img, const float *mask, float *out,
                          int C, int Cm, int H, int W, float v, float l_grad,
                          float l_eucl, int iters) {
  if (C < 1 || Cm < 1 || H < 1 || W < 1) return -1;
  const size_t n = (size_t)H * W;
  const float r2 = sqrtf(2.0f);
  const float ld[9] = {0.f, 0.f, 0.f, r2, 1.0f, r2, 0.f, 0.f, 0.f};
  init_dist(mask, out, (size_t)Cm * n, v);
  float *ia = (float *)malloc((size_t)C * n * sizeof(float));
  float *ib = (float *)malloc((size_t)C * n * sizeof(float));
  float *da = (float *)malloc(n * sizeof(float));
  float *db = (float *)malloc(n * sizeof(float));
  if (!ia || !ib || !da || !db) {
    free(ia); free(ib); free(da); free(db);
    return -2;
  }
  const ptrdiff_t cs = (ptrdiff_t)n;
  memcpy(ia, img, (size_t)C * n * sizeof(float));
  memcpy(da, out, n * sizeof(float));
  for (int it = 0; it < iters; ++it) {
    pass_pair(ia, da, C, cs, H, W, 1, 0, W, 1, ld, l_grad, l_eucl);
    /* (H, W) -> (W, H) */
    permute_copy(ia, ib, C, cs, 1, W, H, 0, 1, W);
    permute_copy(da, db, 1, cs, 1, W, H, 0, 1, W);
    pass_pair(ib, db, C, cs, W, H, 1, 0, H, 1, ld, l_grad, l_eucl);
    /* (W, H) -> (H, W) */
    permute_copy(ib, ia, C, cs, 1, H, W, 0, 1, H);
    permute_copy(db, da, 1, cs, 1, H, W, 0, 1, H);
  }
  memcpy(out, da, n * sizeof(float));
  free(ia); free(ib); free(da); free(db);
  return 0;
}

/* ---- Toivanen's raster scan (accuracy cross-check) -----------------------
 * The reference's alternative CPU transform (geodis_toivanen.cpp:40-147 for
 * 2D, :168-329 for 3D, drivers :149-166 and :331-348): one forward raster
 * scan over the whole grid in memory order, each element relaxed from its
 * causal neighbours (4 in 2D, 13 in 3D) with values already updated in the
 * same scan, then one backward scan with the mirrored neighbour set;
 * `iters` times.  Candidate, channel sum and min rule are those of the
 * raster-scan transform; step lengths are sqrt of the squared spacings of
 * the axes a step moves along, summed in (D, H, W) order from 0 (2D: 1 and
 * sqrtf(2)).  Only mask channel 0 is transformed; out holds fl(v * mask).
 * The neighbour sets of the two scans together are the full 8 / 26
 * neighbourhood, the graph of the raster-scan transform, so both converge to
 * the same values (bit for bit where the step tables agree). */

/* 3D neighbour offsets (dz, dh, dw) of the forward and backward scans, in
 * the reference's visiting order (geodis_toivanen.cpp:185-187, 212-214). */
static const int tv_fwd3[13][3] = {
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {0, -1, -1}, {0, -1, 0},
    {0, -1, 1},   {0, 0, -1},  {1, -1, -1}, {1, -1, 0},  {1, -1, 1}, {1, 0, -1}};
static const int tv_bwd3[13][3] = {
    {-1, 0, 1}, {-1, 1, -1}, {-1, 1, 0}, {-1, 1, 1}, {0, 0, 1}, {0, 1, -1}, {0, 1, 0},
    {0, 1, 1},  {1, 0, 0},   {1, 0, 1},  {1, 1, -1}, {1, 1, 0}, {1, 1, 1}};
/* 2D (dh, dw): geodis_toivanen.cpp:55-56, 102-103 */
static const int tv_fwd2[4][3] = {{0, -1, -1}, {0, -1, 0}, {0, -1, 1}, {0, 0, -1}};
static const int tv_bwd2[4][3] = {{0, 0, 1}, {0, 1, -1}, {0, 1, 0}, {0, 1, 1}};

static void tv_scan(const float *img, float *d, int C, ptrdiff_t cs, int D, int H, int W,
                    const int (*off)[3], const float *ld, int n, int forward,
                    float l_grad, float l_eucl) {
  const ptrdiff_t total = (ptrdiff_t)D * H * W;
  for (ptrdiff_t t = 0; t < total; ++t) {
    const ptrdiff_t p = forward ? t : total - 1 - t;
    const int z = (int)(p / ((ptrdiff_t)H * W));
    const int h = (int)((p / W) % H);
    const int w = (int)(p % W);
    float nd = d[p];
    for (int k = 0; k < n; ++k) {
      const int zq = z + off[k][0], hq = h + off[k][1], wq = w + off[k][2];
      if (zq < 0 || zq >= D || hq < 0 || hq >= H || wq < 0 || wq >= W) continue;
      const ptrdiff_t q = ((ptrdiff_t)zq * H + hq) * W + wq;
      float l;
      if (C == 1) {
        l = fabsf(img[p] - img[q]);
      } else {
        l = 0.0f;
        for (int c = 0; c < C; ++c) l += fabsf(img[c * cs + p] - img[c * cs + q]);
      }
      const float cur = (d[q] + l_eucl * ld[k]) + l_grad * l;
      nd = ref_min(nd, cur);
    }
    d[p] = nd;
  }
}

/* sqrt of the squared spacings of the axes an offset moves along, summed in
 * (D, H, W) order (geodis_toivanen.cpp:189-209, 216-236). */
static void tv_table(const int (*off)[3], int n, const float sq[3], float *ld) {
  for (int k = 0; k < n; ++k) {
    float s = 0.0f;
    for (int a = 0; a < 3; ++a)
      if (off[k][a] != 0) s += sq[a];
    ld[k] = sqrtf(s);
  }
}

int fgo_toivanen2d(const float *img, const float *mask, float *out, int C, int Cm, int H,
                   int W, float v, float l_grad, float l_eucl, int iters) {
  if (C < 1 || Cm < 1 || H < 1 || W < 1) return -1;
  const float r2 = sqrtf(2.0f);
  const float ldf[4] = {r2, 1.0f, r2, 1.0f}, ldb[4] = {1.0f, r2, 1.0f, r2};
  init_dist(mask, out, (size_t)Cm * H * W, v);
  for (int it = 0; it < iters; ++it) {
    tv_scan(img, out, C, (ptrdiff_t)H * W, 1, H, W, tv_fwd2, ldf, 4, 1, l_grad, l_eucl);
    tv_scan(img, out, C, (ptrdiff_t)H * W, 1, H, W, tv_bwd2, ldb, 4, 0, l_grad, l_eucl);
  }
  return 0;
}

int fgo_toivanen3d(const float *img, const float *mask, float *out, int C, int Cm, int D,
                   int H, int W, const float spacing[3], float v, float l_grad, float l_eucl,
                   int iters) {
  if (C < 1 || Cm < 1 || D < 1 || H < 1 || W < 1) return -1;
  float sq[3], ldf[13], ldb[13];
  for (int k = 0; k < 3; ++k) sq[k] = spacing[k] * spacing[k];
  tv_table(tv_fwd3, 13, sq, ldf);
  tv_table(tv_bwd3, 13, sq, ldb);
  init_dist(mask, out, (size_t)Cm * D * H * W, v);
  const ptrdiff_t cs = (ptrdiff_t)D * H * W;
  for (int it = 0; it < iters; ++it) {
    tv_scan(img, out, C, cs, D, H, W, tv_fwd3, ldf, 13, 1, l_grad, l_eucl);
    tv_scan(img, out, C, cs, D, H, W, tv_bwd3, ldb, 13, 0, l_grad, l_eucl);
  }
  return 0;
}
