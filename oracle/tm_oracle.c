/*
 * tm_oracle.c -- CPU restatement of the torchmorph operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled
 * into oracle/libtm_oracle.so and loaded by tests/, by the smoke check and by
 * the CPU-baseline leg of bench.py.  Nothing in the shipped package links,
 * loads or calls it; the product path is the HIP library under
 * torchmorph_amd/csrc and fails loudly when that library is missing.
 *
 * Every routine is a single-item (one (B,C) slice), single-threaded,
 * deliberately plain restatement.  It follows:
 *
 *  - EDT:  scipy.ndimage.distance_transform_edt semantics, which the
 *          reference tests pin (test/test_distance_transform.py:32-69,
 *          :289-320, :397-408).  The feature transform is the separable
 *          lower-envelope algorithm of Maurer, Qi and Raghavan (PAMI 2003),
 *          one pass per axis in axis order 0..n-1, with SciPy's
 *          conventions: double arithmetic, site removal when
 *          c*vR - b*uR - a*wR - a*b*c > 0, query pointer advancing only on a
 *          strictly smaller distance (lower index wins ties), lines with no
 *          site left untouched, and foreground voxels initialised to the
 *          feature (-1, 0, ..., 0) (the all-foreground "virtual background"
 *          the reference emulates at edt_kernel.cu:415-437).  Distances
 *          follow scipy/ndimage/_morphology.py:2582-2598 (int32 deltas,
 *          times sampling, squared, summed in axis order, sqrt, in double).
 *          Pinned against SciPy golden vectors (tests/golden) and a brute
 *          force search (tests/test_oracle_edt.py).
 *  - CDT:  scipy.ndimage.distance_transform_cdt semantics
 *          (_morphology.py:2355-2414): int32 distances initialised to -1 on
 *          foreground and 0 on background, two raster passes (forward, then
 *          over the axis-reversed array) that relax each foreground voxel
 *          from the causal half of the 3^n metric structure in structure
 *          order with a strict improvement test.  The reference returns the
 *          same values as float32 with -1 for unreachable voxels
 *          (cdt_kernel.cu:47-55).  Pinned against SciPy golden vectors.
 *  - grey: the reference gather formulation (grey_kernel.cu:101-170 with
 *          offsets built at :247-292 and border modes at :63-97):
 *          erosion  out[p] = min_i in[p + delta_i] - s_i,
 *          dilation out[p] = max_i in[p - delta_i] + s_i,
 *          delta_i = coord_i - (size/2 + origin).  Pinned against SciPy.
 *  - binary: the reference gather formulation (binary_kernel.cu:56-127,
 *          offsets :199-222): erosion ANDs in[p + delta_i], dilation ORs
 *          in[p - delta_i], out-of-range reads give border_value.
 *  - brute-force DT: the reference's bfdt (bfdt_kernel.cu:23-129, host
 *          :170-296): every foreground voxel takes the background voxel at
 *          the smallest distance, the first one in flat order on ties
 *          (strict minimum), euclidean / taxicab / chessboard with sampling
 *          applied to every metric; background voxels 0 and their own index;
 *          no background: +inf (euclidean) or 4294967295, index 0.  In double
 *          (the reference's float arithmetic is not reproducible); pinned
 *          against scipy.ndimage.distance_transform_bf and against the EDT
 *          oracle (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TMO_MAX_NDIM 8

static int64_t prod(const int64_t* shape, int lo, int hi) {
    int64_t p = 1;
    for (int i = lo; i < hi; ++i) p *= shape[i];
    return p;
}

/* ------------------------------------------------------------------ EDT -- */

/* One lower-envelope pass along axis d for one line.
 * feat: ndim planes of int32, plane stride n; base: flat index of the line's
 * first voxel; step: flat stride along axis d; coor: the line's coordinates
 * on the other axes (coor[d] unused). */
static void edt_line(int32_t* feat, int64_t n, int ndim, int d, int64_t base, int64_t step,
                     int64_t len, const int64_t* coor, const double* s,
                     int32_t* f /* len*ndim scratch */, int64_t* g /* len */) {
    for (int64_t i = 0; i < len; ++i)
        for (int j = 0; j < ndim; ++j) f[i * ndim + j] = feat[j * n + base + i * step];

    int64_t l = -1;
    for (int64_t i = 0; i < len; ++i) {
        if (f[i * ndim + 0] < 0) continue;
        double wR = 0.0;
        for (int j = 0; j < ndim; ++j) {
            if (j == d) continue;
            double t = (double)(f[i * ndim + j] - coor[j]) * s[j];
            wR += t * t;
        }
        while (l >= 1) {
            int64_t v = g[l], u = g[l - 1];
            double a = (double)(f[v * ndim + d] - f[u * ndim + d]) * s[d];
            double b = (double)(f[i * ndim + d] - f[v * ndim + d]) * s[d];
            double c = a + b;
            double uR = 0.0, vR = 0.0;
            for (int j = 0; j < ndim; ++j) {
                if (j == d) continue;
                double tu = (double)(f[u * ndim + j] - coor[j]) * s[j];
                double tv = (double)(f[v * ndim + j] - coor[j]) * s[j];
                uR += tu * tu;
                vR += tv * tv;
            }
            if (c * vR - b * uR - a * wR - a * b * c <= 0.0) break;
            --l;
        }
        g[++l] = i;
    }
    if (l < 0) return; /* no site on this line: features stay as they are */
    int64_t top = l;
    l = 0;
    for (int64_t i = 0; i < len; ++i) {
        double d1 = 0.0;
        for (int j = 0; j < ndim; ++j) {
            double t = (double)(f[g[l] * ndim + j] - (j == d ? i : coor[j])) * s[j];
            d1 += t * t;
        }
        while (l < top) {
            double d2 = 0.0;
            for (int j = 0; j < ndim; ++j) {
                double t = (double)(f[g[l + 1] * ndim + j] - (j == d ? i : coor[j])) * s[j];
                d2 += t * t;
            }
            if (d1 <= d2) break;
            d1 = d2;
            ++l;
        }
        for (int j = 0; j < ndim; ++j) feat[j * n + base + i * step] = f[g[l] * ndim + j];
    }
}

/* Exact EDT of one item.  mask: nonzero = foreground.  sampling may be NULL
 * (unit).  features: ndim planes of n int32 (may be NULL).  dist: n doubles
 * (may be NULL).  Returns 0 on success. */
int tmo_edt(const uint8_t* mask, int ndim, const int64_t* shape, const double* sampling,
            int32_t* features, double* dist) {
    if (ndim < 1 || ndim > TMO_MAX_NDIM) return 1;
    int64_t n = prod(shape, 0, ndim);
    double s[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) s[j] = sampling ? sampling[j] : 1.0;
    int32_t* feat = features ? features : (int32_t*)malloc(sizeof(int32_t) * n * ndim);
    int64_t maxlen = 1;
    for (int j = 0; j < ndim; ++j) maxlen = shape[j] > maxlen ? shape[j] : maxlen;
    int32_t* f = (int32_t*)malloc(sizeof(int32_t) * maxlen * ndim);
    int64_t* g = (int64_t*)malloc(sizeof(int64_t) * maxlen);
    if (!feat || !f || !g) return 2;

    /* initial features: background -> own coordinates, foreground -> (-1,0..) */
    for (int64_t p = 0; p < n; ++p) {
        int64_t rem = p;
        for (int j = ndim - 1; j >= 0; --j) {
            int64_t c = rem % shape[j];
            rem /= shape[j];
            feat[j * n + p] = mask[p] ? 0 : (int32_t)c;
        }
        if (mask[p]) feat[0 * n + p] = -1;
    }
    int64_t coor[TMO_MAX_NDIM];
    for (int d = 0; d < ndim; ++d) {
        int64_t step = prod(shape, d + 1, ndim);
        int64_t outer = prod(shape, 0, d);
        for (int64_t o = 0; o < outer; ++o) {
            for (int64_t in = 0; in < step; ++in) {
                int64_t rem = o;
                for (int j = d - 1; j >= 0; --j) { coor[j] = rem % shape[j]; rem /= shape[j]; }
                rem = in;
                for (int j = ndim - 1; j > d; --j) { coor[j] = rem % shape[j]; rem /= shape[j]; }
                coor[d] = 0;
                int64_t base = o * shape[d] * step + in;
                edt_line(feat, n, ndim, d, base, step, shape[d], coor, s, f, g);
            }
        }
    }
    if (dist) {
        for (int64_t p = 0; p < n; ++p) {
            int64_t rem = p;
            int64_t c[TMO_MAX_NDIM];
            for (int j = ndim - 1; j >= 0; --j) { c[j] = rem % shape[j]; rem /= shape[j]; }
            double acc = 0.0;
            for (int j = 0; j < ndim; ++j) {
                int32_t delta = feat[j * n + p] - (int32_t)c[j];
                double t = (double)delta * s[j];
                t = t * t;
                acc = (j == 0) ? t : acc + t;
            }
            dist[p] = sqrt(acc);
        }
    }
    free(f);
    free(g);
    if (!features) free(feat);
    return 0;
}

/* ------------------------------------------------------------------ CDT -- */

static void cdt_pass(int32_t* dt, int32_t* ft, int ndim, const int64_t* shape,
                     const int64_t* offs, const int64_t (*rel)[TMO_MAX_NDIM], int noff) {
    int64_t n = prod(shape, 0, ndim);
    int64_t c[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) c[j] = 0;
    for (int64_t p = 0; p < n; ++p) {
        int32_t value = dt[p];
        if (value != 0) {
            int32_t best = value;
            int64_t src = p;
            for (int k = 0; k < noff; ++k) {
                int inb = 1;
                for (int j = 0; j < ndim; ++j) {
                    int64_t q = c[j] + rel[k][j];
                    if (q < 0 || q >= shape[j]) { inb = 0; break; }
                }
                if (!inb) continue;
                int32_t tt = dt[p + offs[k]];
                if (tt >= 0 && (best < 0 || tt + 1 < best)) {
                    best = tt + 1;
                    src = p + offs[k];
                }
            }
            dt[p] = best;
            if (ft) ft[p] = ft[src];
        }
        for (int j = ndim - 1; j >= 0; --j) {
            if (++c[j] < shape[j]) break;
            c[j] = 0;
        }
    }
}

static void reverse_all(int32_t* a, int64_t n) {
    for (int64_t i = 0, k = n - 1; i < k; ++i, --k) {
        int32_t t = a[i];
        a[i] = a[k];
        a[k] = t;
    }
}

/* Chamfer DT of one item.  metric: 0 = chessboard (full 3^n structure),
 * 1 = taxicab (cross).  dist: n int32 (-1 where unreachable).  features:
 * ndim planes of n int32 coordinates (may be NULL). */
int tmo_cdt(const uint8_t* mask, int ndim, const int64_t* shape, int metric, int32_t* dist,
            int32_t* features) {
    if (ndim < 1 || ndim > TMO_MAX_NDIM) return 1;
    int64_t n = prod(shape, 0, ndim);
    int64_t stride[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) stride[j] = prod(shape, j + 1, ndim);
    /* causal half of the 3^n structure, in structure raster order */
    int64_t offs[3280];
    int64_t rel[3280][TMO_MAX_NDIM];
    int noff = 0;
    int64_t total = 1;
    for (int j = 0; j < ndim; ++j) total *= 3;
    for (int64_t e = 0; e < total / 2; ++e) {
        int64_t rem = e, nz = 0, off = 0;
        int64_t r[TMO_MAX_NDIM];
        for (int j = ndim - 1; j >= 0; --j) { r[j] = rem % 3 - 1; rem /= 3; }
        for (int j = 0; j < ndim; ++j) { nz += r[j] != 0; off += r[j] * stride[j]; }
        if (metric == 1 && nz > 1) continue;
        offs[noff] = off;
        for (int j = 0; j < ndim; ++j) rel[noff][j] = r[j];
        ++noff;
    }
    int32_t* ft = features ? (int32_t*)malloc(sizeof(int32_t) * n) : NULL;
    for (int64_t p = 0; p < n; ++p) {
        dist[p] = mask[p] ? -1 : 0;
        if (ft) ft[p] = (int32_t)p;
    }
    cdt_pass(dist, ft, ndim, shape, offs, (const int64_t(*)[TMO_MAX_NDIM])rel, noff);
    reverse_all(dist, n);
    if (ft) reverse_all(ft, n);
    cdt_pass(dist, ft, ndim, shape, offs, (const int64_t(*)[TMO_MAX_NDIM])rel, noff);
    reverse_all(dist, n);
    if (ft) {
        reverse_all(ft, n);
        for (int64_t p = 0; p < n; ++p) {
            int64_t rem = ft[p];
            for (int j = ndim - 1; j >= 0; --j) {
                features[j * n + p] = (int32_t)(rem % shape[j]);
                rem /= shape[j];
            }
        }
        free(ft);
    }
    return 0;
}

/* ------------------------------------------------------- border handling -- */

static int64_t map_coord(int64_t c, int64_t size, int mode, int* use_cval) {
    *use_cval = 0;
    if (c >= 0 && c < size) return c;
    switch (mode) {
    case 0: *use_cval = 1; return 0; /* constant */
    case 2: return c < 0 ? 0 : size - 1; /* nearest */
    case 1: { /* reflect: half-sample symmetric */
        int64_t period = 2 * size;
        int64_t m = ((c % period) + period) % period;
        return m >= size ? period - 1 - m : m;
    }
    case 3: { /* mirror: whole-sample symmetric */
        int64_t period = 2 * (size - 1);
        if (period <= 0) return 0;
        int64_t m = ((c % period) + period) % period;
        return m >= size ? period - m : m;
    }
    case 4: return ((c % size) + size) % size; /* wrap */
    default: *use_cval = 1; return 0;
    }
}

/* ----------------------------------------------------------------- grey -- */

/* min / max as the GPU's v_min_f32 / v_max_f32 evaluate them (measured on
 * MI355X): a NaN operand yields the other operand, and -0 orders below +0
 * whichever operand comes first, so a window holding both zeros erodes to -0
 * and dilates to +0 in any scan order. */
static inline float tmo_fmax(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  if (a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}
static inline float tmo_fmin(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  if (a == b) return signbit(a) ? a : b;
  return a < b ? a : b;
}

int tmo_grey(const float* in, int ndim, const int64_t* shape, const float* structure,
             const uint8_t* footprint, const int64_t* se_shape, const int64_t* origin, int mode,
             float cval, int dilation, float* out) {
    if (ndim < 1 || ndim > TMO_MAX_NDIM) return 1;
    int64_t n = prod(shape, 0, ndim);
    int64_t ne = prod(se_shape, 0, ndim);
    int64_t* delta = (int64_t*)malloc(sizeof(int64_t) * ne * ndim);
    float* sv = (float*)malloc(sizeof(float) * ne);
    int64_t na = 0;
    for (int64_t e = 0; e < ne; ++e) {
        if (footprint && !footprint[e]) continue;
        int64_t rem = e;
        int64_t r[TMO_MAX_NDIM];
        for (int j = ndim - 1; j >= 0; --j) { r[j] = rem % se_shape[j]; rem /= se_shape[j]; }
        for (int j = 0; j < ndim; ++j) {
            int64_t dd = r[j] - (se_shape[j] / 2 + origin[j]);
            delta[na * ndim + j] = dilation ? -dd : dd;
        }
        sv[na] = structure ? structure[e] : 0.0f;
        ++na;
    }
    int64_t c[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) c[j] = 0;
    for (int64_t p = 0; p < n; ++p) {
        float acc = dilation ? -INFINITY : INFINITY;
        for (int64_t k = 0; k < na; ++k) {
            int64_t flat = 0;
            int uc = 0;
            for (int j = 0; j < ndim && !uc; ++j) {
                int64_t m = map_coord(c[j] + delta[k * ndim + j], shape[j], mode, &uc);
                flat = flat * shape[j] + m;
            }
            float v = uc ? cval : in[flat];
            if (dilation) acc = tmo_fmax(acc, v + sv[k]);
            else acc = tmo_fmin(acc, v - sv[k]);
        }
        out[p] = acc;
        for (int j = ndim - 1; j >= 0; --j) {
            if (++c[j] < shape[j]) break;
            c[j] = 0;
        }
    }
    free(delta);
    free(sv);
    return 0;
}

/* --------------------------------------------------------------- binary -- */

int tmo_binary(const uint8_t* in, int ndim, const int64_t* shape, const uint8_t* structure,
               const int64_t* se_shape, const int64_t* origin, int border_value, int dilation,
               uint8_t* out) {
    if (ndim < 1 || ndim > TMO_MAX_NDIM) return 1;
    int64_t n = prod(shape, 0, ndim);
    int64_t ne = prod(se_shape, 0, ndim);
    int64_t* delta = (int64_t*)malloc(sizeof(int64_t) * (ne ? ne : 1) * ndim);
    int64_t na = 0;
    for (int64_t e = 0; e < ne; ++e) {
        if (!structure[e]) continue;
        int64_t rem = e;
        int64_t r[TMO_MAX_NDIM];
        for (int j = ndim - 1; j >= 0; --j) { r[j] = rem % se_shape[j]; rem /= se_shape[j]; }
        for (int j = 0; j < ndim; ++j) {
            int64_t dd = r[j] - (se_shape[j] / 2 + origin[j]);
            delta[na * ndim + j] = dilation ? -dd : dd;
        }
        ++na;
    }
    int64_t c[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) c[j] = 0;
    for (int64_t p = 0; p < n; ++p) {
        int acc = dilation ? 0 : 1;
        if (na == 0) acc = dilation ? 0 : 1;
        for (int64_t k = 0; k < na; ++k) {
            int64_t flat = 0;
            int inb = 1;
            for (int j = 0; j < ndim; ++j) {
                int64_t q = c[j] + delta[k * ndim + j];
                if (q < 0 || q >= shape[j]) { inb = 0; break; }
                flat = flat * shape[j] + q;
            }
            int v = inb ? (in[flat] != 0) : (border_value != 0);
            if (dilation) { if (v) { acc = 1; break; } }
            else if (!v) { acc = 0; break; }
        }
        out[p] = (uint8_t)acc;
        for (int j = ndim - 1; j >= 0; --j) {
            if (++c[j] < shape[j]) break;
            c[j] = 0;
        }
    }
    free(delta);
    return 0;
}

/* Brute-force distance transform of one item (metric 0 euclidean, 1
 * taxicab, 2 chessboard).  dist: distances (sqrt taken for euclidean);
 * features: ndim planes of the chosen background's coordinates. */
int tmo_bfdt(const uint8_t* mask, int ndim, const int64_t* shape, const double* sampling,
             int metric, double* dist, int32_t* features) {
    if (ndim < 1 || ndim > TMO_MAX_NDIM || metric < 0 || metric > 2) return 1;
    const int64_t n = prod(shape, 0, ndim);
    double s[TMO_MAX_NDIM];
    for (int j = 0; j < ndim; ++j) s[j] = sampling ? sampling[j] : 1.0;
    int32_t* coords = (int32_t*)malloc(sizeof(int32_t) * n * ndim);
    if (!coords) return 2;
    for (int64_t p = 0; p < n; ++p) {
        int64_t rem = p;
        for (int j = ndim - 1; j >= 0; --j) {
            coords[p * ndim + j] = (int32_t)(rem % shape[j]);
            rem /= shape[j];
        }
    }
    for (int64_t p = 0; p < n; ++p) {
        const int32_t* c = coords + p * ndim;
        int64_t best_q = -1;
        double best = 0.0;
        if (mask[p]) {
            for (int64_t q = 0; q < n; ++q) {
                if (mask[q]) continue;
                const int32_t* y = coords + q * ndim;
                double acc = 0.0;
                for (int j = 0; j < ndim; ++j) {
                    double t = (double)(c[j] - y[j]) * s[j];
                    if (metric == 0) t = t * t;
                    else if (t < 0) t = -t;
                    if (j == 0) acc = t;
                    else if (metric == 2) acc = t > acc ? t : acc;
                    else acc = acc + t;
                }
                if (best_q < 0 || acc < best) { best = acc; best_q = q; }
            }
        }
        double out;
        const int32_t* f;
        int32_t zero[TMO_MAX_NDIM] = {0};
        if (!mask[p]) { out = 0.0; f = c; }
        else if (best_q < 0) { out = metric == 0 ? INFINITY : 4294967295.0; f = zero; }
        else { out = metric == 0 ? sqrt(best) : best; f = coords + best_q * ndim; }
        if (dist) dist[p] = out;
        if (features)
            for (int j = 0; j < ndim; ++j) features[j * n + p] = f[j];
    }
    free(coords);
    return 0;
}
