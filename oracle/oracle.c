/*
 * oracle.c -- CPU restatement of the reference's BCSR x dense path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Build: oracle/Makefile (gcc, -O2, no fast-math: the fp32 products
 * must keep the reference's summation order).
 *
 * Every function cites the reference file:line it restates; oracle.h lists
 * what pins each one.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * std::mt19937 (matrix_generator.cpp:34 `std::mt19937 rng(seed)`), the
 * 32-bit Mersenne Twister with the C++11 parameters.
 * ------------------------------------------------------------------------ */
void orc_mt_seed(orc_mt *g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i) {
        uint32_t prev = g->mt[i - 1];
        g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
    }
    g->idx = 624;
}

static void orc_mt_twist(orc_mt *g) {
    for (int i = 0; i < 624; ++i) {
        uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        g->mt[i] = v;
    }
    g->idx = 0;
}

uint32_t orc_mt_next(orc_mt *g) {
    if (g->idx >= 624) orc_mt_twist(g);
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orc_mt_stream(uint32_t seed, uint32_t *out, size_t n) {
    orc_mt g;
    orc_mt_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) out[i] = orc_mt_next(&g);
}

/*
 * generate_canonical<double, 53>(mt19937) as libstdc++ computes it: two
 * 32-bit draws, sum = x0 + x1 * 2^32 rounded once to double, divided by 2^64,
 * and a result of exactly 1.0 replaced by nextafter(1, 0).
 */
double orc_canonical(orc_mt *g) {
    double sum = (double)orc_mt_next(g);
    sum += (double)orc_mt_next(g) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* bernoulli_distribution(p): one canonical draw, true iff draw < p
 * (matrix_generator.cpp:50, 89, 102, 121). */
int orc_bernoulli(orc_mt *g, double p) { return orc_canonical(g) < p; }

/* uniform_real_distribution<double>(a, b): canonical * (b - a) + a
 * (matrix_generator.cpp:20, 58). */
double orc_uniform(orc_mt *g, double a, double b) { return orc_canonical(g) * (b - a) + a; }

/*
 * uniform_int_distribution<I>(a, b) for a 32-bit engine and a range below
 * 2^32: Lemire's multiply-shift with rejection of the biased low words
 * (matrix_generator.cpp:15, 54).
 */
int64_t orc_uniform_int(orc_mt *g, int64_t a, int64_t b) {
    uint32_t range = (uint32_t)(b - a) + 1u;
    uint64_t product = (uint64_t)orc_mt_next(g) * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        uint32_t threshold = (uint32_t)(0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)orc_mt_next(g) * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return (int64_t)(product >> 32) + a;
}

/* ---------------------------------------------------------------------------
 * IEEE binary32 <-> binary16, round to nearest even (__float2half /
 * __half2float as used at nvbench_main.cu:72 and Matrix.cu:30).
 * ------------------------------------------------------------------------ */
uint16_t orc_f2h(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {
        if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
        return (uint16_t)(sign | 0x7c00u);
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 -> inf */
    if (ax < 0x38800000u) {                                    /* below 2^-14 */
        if (ax < 0x33000000u) return (uint16_t)sign;            /* below 2^-25 */
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;
        uint32_t r = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (r & 1u))) r++;
        return (uint16_t)(sign | r);
    }
    uint32_t r = (ax - 0x38000000u) >> 13;
    uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r++;
    return (uint16_t)(sign | r);
}

float orc_h2f(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {
            e = 113;
            while (!(m & 0x400u)) {
                m <<= 1;
                e--;
            }
            m &= 0x3ffu;
            x = sign | (e << 23) | (m << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void orc_f32_to_f16(const float *in, uint16_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_f2h(in[i]);
}

void orc_f16_to_f32(const uint16_t *in, float *out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_h2f(in[i]);
}

/* ---------------------------------------------------------------------------
 * mg::generate_matrix<T> (matrix_generator.cpp:28-140).
 * ------------------------------------------------------------------------ */
typedef struct {
    orc_mt g;
    int elem_type;  /* ORC_T_* */
    int value_dist; /* ORC_VAL_* */
    double lo, hi;
} orc_gen;

/* One value of generate_dense_block<T> (matrix_generator.cpp:11-26): ints use
 * uniform_int_distribution<long long>(-1000, 1000), reals U(-1, 1) narrowed
 * to T.  ORC_VAL_LOGUNIFORM (an extension) draws a sign, then a magnitude
 * lo * (hi/lo)^u. */
static double orc_block_value(orc_gen *s) {
    if (s->elem_type == ORC_T_INT || s->elem_type == ORC_T_LONGLONG)
        return (double)orc_uniform_int(&s->g, -1000, 1000);
    double v;
    if (s->value_dist == ORC_VAL_LOGUNIFORM) {
        double sgn = orc_canonical(&s->g) < 0.5 ? -1.0 : 1.0;
        double u = orc_canonical(&s->g);
        v = sgn * (s->lo * pow(s->hi / s->lo, u));
    } else {
        v = orc_uniform(&s->g, -1.0, 1.0);
    }
    if (s->elem_type == ORC_T_FLOAT) v = (double)(float)v;
    return v;
}

/* Per-element value of the "random" pattern (matrix_generator.cpp:52-60):
 * ints use uniform_int_distribution<int>, which draws identically. */
static double orc_elem_value(orc_gen *s) { return orc_block_value(s); }

/* fill_block (matrix_generator.cpp:37-46): a dense block of the clipped
 * extent, drawn row-major. */
static void orc_fill_block(orc_gen *s, double *out, int rows, int cols, int r0, int c0,
                           int bh, int bw) {
    int r1 = r0 + bh < rows ? r0 + bh : rows;
    int c1 = c0 + bw < cols ? c0 + bw : cols;
    for (int i = r0; i < r1; ++i)
        for (int j = c0; j < c1; ++j) out[(size_t)i * cols + j] = orc_block_value(s);
}

/* Like orc_fill_block, each element kept with probability rho (extension). */
static void orc_fill_block_rho(orc_gen *s, double *out, int rows, int cols, int r0, int c0,
                               int bh, int bw, double rho) {
    int r1 = r0 + bh < rows ? r0 + bh : rows;
    int c1 = c0 + bw < cols ? c0 + bw : cols;
    for (int i = r0; i < r1; ++i)
        for (int j = c0; j < c1; ++j)
            out[(size_t)i * cols + j] = orc_bernoulli(&s->g, rho) ? orc_block_value(s) : 0.0;
}

static int orc_generate_core(orc_gen *s, int rows, int cols, double sparsity,
                             const char *pattern, int bh, int bw, int band_w, double rho,
                             double *out) {
    memset(out, 0, sizeof(double) * (size_t)rows * (size_t)cols);
    if (strcmp(pattern, "random") == 0) {
        /* matrix_generator.cpp:49-66 */
        double p = 1.0 - sparsity;
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j)
                out[(size_t)i * cols + j] = orc_bernoulli(&s->g, p) ? orc_elem_value(s) : 0.0;
    } else if (strcmp(pattern, "checkerboard") == 0) {
        /* matrix_generator.cpp:68-74 */
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bh - 1) / bh;
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (((i % 2) ^ (j % 2)) == 0) orc_fill_block(s, out, rows, cols, i * bh, j * bh, bh, bh);
    } else if (strcmp(pattern, "diagonal") == 0) {
        /* matrix_generator.cpp:76-81 */
        int m = rows < cols ? rows : cols;
        for (int i = 0; i < m; ++i) out[(size_t)i * cols + i] = orc_block_value(s);
    } else if (strcmp(pattern, "blockdiagonal") == 0) {
        /* matrix_generator.cpp:83-85: min(r,c)/bs + 1 diagonal blocks */
        int m = rows < cols ? rows : cols;
        int n = m / bh + 1;
        for (int i = 0; i < n; ++i) orc_fill_block(s, out, rows, cols, i * bh, i * bh, bh, bh);
    } else if (strcmp(pattern, "blockrandom") == 0) {
        /* matrix_generator.cpp:87-93 (square bh x bh in the reference; the
         * extension allows bh x bw) */
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        double p = 1.0 - sparsity;
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (orc_bernoulli(&s->g, p)) orc_fill_block(s, out, rows, cols, i * bh, j * bw, bh, bw);
    } else if (strcmp(pattern, "pattern64by16") == 0 || strcmp(pattern, "pattern32by16") == 0) {
        /* matrix_generator.cpp:95-134 */
        int brow = pattern[7] == '6' ? 64 : 32, bcol = 16;
        int nbi = (rows + brow - 1) / brow, nbj = (cols + bcol - 1) / bcol;
        double p = 1.0 - sparsity;
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (orc_bernoulli(&s->g, p)) orc_fill_block(s, out, rows, cols, i * brow, j * bcol, brow, bcol);
    } else if (strcmp(pattern, "blockband") == 0) {
        /* extension: block (bi, bj) kept iff bj lies within band_w block-columns
         * of the block-columns that block-row bi's diagonal crosses */
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        for (int i = 0; i < nbi; ++i) {
            int d0 = (i * bh) / bw, d1 = (i * bh + bh - 1) / bw;
            for (int j = 0; j < nbj; ++j)
                if (j >= d0 - band_w && j <= d1 + band_w)
                    orc_fill_block(s, out, rows, cols, i * bh, j * bw, bh, bw);
        }
    } else if (strcmp(pattern, "blockrandom_rho") == 0) {
        /* extension: blockrandom whose kept blocks hold each entry with
         * probability rho (zeros inside a kept block stay stored) */
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        double p = 1.0 - sparsity;
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (orc_bernoulli(&s->g, p))
                    orc_fill_block_rho(s, out, rows, cols, i * bh, j * bw, bh, bw, rho);
    } else {
        return -1;
    }
    return 0;
}

int orc_generate_matrix(int rows, int cols, double sparsity, const char *pattern,
                        int blocksize, uint32_t seed, int elem_type, double *out) {
    orc_gen s;
    orc_mt_seed(&s.g, seed);
    s.elem_type = elem_type;
    s.value_dist = ORC_VAL_UNIFORM;
    s.lo = s.hi = 0.0;
    return orc_generate_core(&s, rows, cols, sparsity, pattern, blocksize, blocksize, 0, 1.0, out);
}

int orc_generate_ex(int rows, int cols, double sparsity, const char *pattern, int bh, int bw,
                    uint32_t seed, int value_dist, double lo, double hi, int band_w, double rho,
                    float *out) {
    size_t n = (size_t)rows * (size_t)cols;
    double *tmp = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!tmp) return -2;
    orc_gen s;
    orc_mt_seed(&s.g, seed);
    s.elem_type = ORC_T_FLOAT;
    s.value_dist = value_dist;
    s.lo = lo;
    s.hi = hi;
    int rc = orc_generate_core(&s, rows, cols, sparsity, pattern, bh, bw, band_w, rho, tmp);
    if (rc == 0)
        for (size_t i = 0; i < n; ++i) out[i] = (float)tmp[i];
    free(tmp);
    return rc;
}

/* ---------------------------------------------------------------------------
 * BCSRMatrix (formats/BCSRMatrix.cu:22-74).  A block is stored iff
 * blockDensity > 0 (miscutil.cpp:76-87: some entry tests true as a half, so
 * -0 counts as zero), block columns ascend within a block-row and each block
 * is row-major with leading dimension bw.
 * ------------------------------------------------------------------------ */
static int orc_block_nonzero(const uint16_t *A, int cols, int r0, int c0, int bh, int bw) {
    for (int x = 0; x < bh; ++x)
        for (int y = 0; y < bw; ++y)
            if (A[(size_t)(r0 + x) * cols + c0 + y] & 0x7fffu) return 1;
    return 0;
}

int64_t orc_bcsr_count(const uint16_t *A, int rows, int cols, int bh, int bw, int32_t *hdr) {
    int nbr = rows / bh, nbc = cols / bw;
    hdr[0] = 0;
    for (int i = 0; i < nbr; ++i) {
        hdr[i + 1] = hdr[i];
        for (int j = 0; j < nbc; ++j)
            if (orc_block_nonzero(A, cols, i * bh, j * bw, bh, bw)) hdr[i + 1]++;
    }
    return hdr[nbr];
}

void orc_bcsr_fill(const uint16_t *A, int rows, int cols, int bh, int bw, const int32_t *hdr,
                   int32_t *idx, uint16_t *data) {
    int nbr = rows / bh, nbc = cols / bw;
    int64_t k = 0;
    (void)hdr;
    for (int i = 0; i < nbr; ++i)
        for (int j = 0; j < nbc; ++j)
            if (orc_block_nonzero(A, cols, i * bh, j * bw, bh, bw)) {
                idx[k] = j;
                for (int x = 0; x < bh; ++x)
                    for (int y = 0; y < bw; ++y)
                        data[k * bh * bw + x * bw + y] = A[(size_t)(i * bh + x) * cols + j * bw + y];
                k++;
            }
}

void orc_bcsr_fill_f32(const float *A, int rows, int cols, int bh, int bw, const int32_t *hdr,
                       int32_t *idx, float *data) {
    int nbr = rows / bh;
    for (int i = 0; i < nbr; ++i)
        for (int64_t k = hdr[i]; k < hdr[i + 1]; ++k) {
            int j = idx[k];
            for (int x = 0; x < bh; ++x)
                for (int y = 0; y < bw; ++y)
                    data[k * bh * bw + x * bw + y] = A[(size_t)(i * bh + x) * cols + j * bw + y];
        }
}

void orc_bcsr_densify(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int block_rows,
                      int block_cols, int bh, int bw, uint16_t *out) {
    size_t cols = (size_t)block_cols * bw;
    memset(out, 0, sizeof(uint16_t) * (size_t)block_rows * bh * cols);
    for (int i = 0; i < block_rows; ++i)
        for (int64_t k = hdr[i]; k < hdr[i + 1]; ++k)
            for (int x = 0; x < bh; ++x)
                for (int y = 0; y < bw; ++y)
                    out[(size_t)(i * bh + x) * cols + (size_t)idx[k] * bw + y] =
                        data[k * bh * bw + x * bw + y];
}

/* CSRMatrix (formats/CSRMatrix.cu:13-40). */
int64_t orc_csr_count(const uint16_t *A, int rows, int cols, int32_t *hdr) {
    hdr[0] = 0;
    for (int i = 0; i < rows; ++i) {
        hdr[i + 1] = hdr[i];
        for (int j = 0; j < cols; ++j)
            if (A[(size_t)i * cols + j] & 0x7fffu) hdr[i + 1]++;
    }
    return hdr[rows];
}

void orc_csr_fill(const uint16_t *A, int rows, int cols, const int32_t *hdr, int32_t *idx,
                  uint16_t *data) {
    int64_t k = 0;
    (void)hdr;
    for (size_t i = 0; i < (size_t)rows * cols; ++i)
        if (A[i] & 0x7fffu) {
            idx[k] = (int32_t)(i % (size_t)cols);
            data[k] = A[i];
            k++;
        }
}

/* ---------------------------------------------------------------------------
 * cpu_matmul_ref (nvbench_main.cu:78-91): for i, for k ascending,
 * a = half2float(A[i,k]); for j, C[i,j] += a * half2float(B[k,j]), fp32 from 0.
 * ------------------------------------------------------------------------ */
void orc_matmul_dense_ref(const uint16_t *A, const uint16_t *B, int M, int K, int N, float *C) {
    float *Bf = (float *)malloc(sizeof(float) * (size_t)K * N);
    orc_f16_to_f32(B, Bf, (size_t)K * N);
    memset(C, 0, sizeof(float) * (size_t)M * N);
    for (int i = 0; i < M; ++i) {
        float *c = C + (size_t)i * N;
        for (int k = 0; k < K; ++k) {
            float a = orc_h2f(A[(size_t)i * K + k]);
            const float *b = Bf + (size_t)k * N;
            for (int j = 0; j < N; ++j) c[j] += a * b[j];
        }
    }
    free(Bf);
}

/* The same sum visited block by block: every k of a stored block in
 * ascending order, absent blocks skipped (they would add +0). */
typedef struct {
    const int32_t *hdr, *idx;
    const uint16_t *data16;
    const float *data32;
    int bh, bw, br0, br1, N;
    const float *Bf;
    float *C;
    double *C64, *Cabs;
} orc_spmm_job;

static void orc_spmm_f32_rows(const orc_spmm_job *jb) {
    int bh = jb->bh, bw = jb->bw, N = jb->N;
    float *arow = (float *)malloc(sizeof(float) * (size_t)bw);
    for (int br = jb->br0; br < jb->br1; ++br) {
        for (int x = 0; x < bh; ++x) {
            float *c = jb->C + ((size_t)br * bh + x) * N;
            memset(c, 0, sizeof(float) * (size_t)N);
            for (int64_t k = jb->hdr[br]; k < jb->hdr[br + 1]; ++k) {
                const uint16_t *blk = jb->data16 + (size_t)k * bh * bw + (size_t)x * bw;
                for (int y = 0; y < bw; ++y) arow[y] = orc_h2f(blk[y]);
                const float *bbase = jb->Bf + (size_t)jb->idx[k] * bw * N;
                for (int y = 0; y < bw; ++y) {
                    float a = arow[y];
                    const float *b = bbase + (size_t)y * N;
                    for (int j = 0; j < N; ++j) c[j] += a * b[j];
                }
            }
        }
    }
    free(arow);
}

static void orc_spmm_f64_rows(const orc_spmm_job *jb) {
    int bh = jb->bh, bw = jb->bw, N = jb->N;
    for (int br = jb->br0; br < jb->br1; ++br) {
        for (int x = 0; x < bh; ++x) {
            size_t row = (size_t)br * bh + x;
            double *c = jb->C64 + row * N;
            double *ca = jb->Cabs ? jb->Cabs + row * N : NULL;
            memset(c, 0, sizeof(double) * (size_t)N);
            if (ca) memset(ca, 0, sizeof(double) * (size_t)N);
            for (int64_t k = jb->hdr[br]; k < jb->hdr[br + 1]; ++k) {
                const float *blk = jb->data32 + (size_t)k * bh * bw + (size_t)x * bw;
                const float *bbase = jb->Bf + (size_t)jb->idx[k] * bw * N;
                for (int y = 0; y < bw; ++y) {
                    double a = blk[y];
                    const float *b = bbase + (size_t)y * N;
                    for (int j = 0; j < N; ++j) c[j] += a * (double)b[j];
                    if (ca) {
                        double aa = fabs(a);
                        for (int j = 0; j < N; ++j) ca[j] += aa * fabs((double)b[j]);
                    }
                }
            }
        }
    }
}

static void *orc_spmm_f32_thread(void *p) {
    orc_spmm_f32_rows((const orc_spmm_job *)p);
    return NULL;
}

static void *orc_spmm_f64_thread(void *p) {
    orc_spmm_f64_rows((const orc_spmm_job *)p);
    return NULL;
}

static void orc_run_split(orc_spmm_job base, int nthreads, void *(*fn)(void *)) {
    int nrows = base.br1 - base.br0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nrows) nthreads = nrows > 0 ? nrows : 1;
    if (nthreads == 1) {
        fn(&base);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    orc_spmm_job *jobs = (orc_spmm_job *)malloc(sizeof(orc_spmm_job) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = base;
        jobs[t].br0 = base.br0 + (int)((int64_t)nrows * t / nthreads);
        jobs[t].br1 = base.br0 + (int)((int64_t)nrows * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, fn, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

void orc_bcsr_spmm_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int bh,
                       int bw, int br0, int br1, const float *Bf, int N, float *C, int nthreads) {
    orc_spmm_job jb;
    memset(&jb, 0, sizeof(jb));
    jb.hdr = hdr;
    jb.idx = idx;
    jb.data16 = data;
    jb.bh = bh;
    jb.bw = bw;
    jb.br0 = br0;
    jb.br1 = br1;
    jb.N = N;
    jb.Bf = Bf;
    jb.C = C;
    orc_run_split(jb, nthreads, orc_spmm_f32_thread);
}

void orc_bcsr_spmm_f64(const int32_t *hdr, const int32_t *idx, const float *data, int bh, int bw,
                       int br0, int br1, const float *Bf, int N, double *C, double *Cabs,
                       int nthreads) {
    orc_spmm_job jb;
    memset(&jb, 0, sizeof(jb));
    jb.hdr = hdr;
    jb.idx = idx;
    jb.data32 = data;
    jb.bh = bh;
    jb.bw = bw;
    jb.br0 = br0;
    jb.br1 = br1;
    jb.N = N;
    jb.Bf = Bf;
    jb.C64 = C;
    jb.Cabs = Cabs;
    orc_run_split(jb, nthreads, orc_spmm_f64_thread);
}

void orc_csr_spmm_acc_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int M, const float *Bf,
                          int N, float *C) {
    for (int i = 0; i < M; ++i) {
        float *c = C + (size_t)i * N;
        for (int32_t p = hdr[i]; p < hdr[i + 1]; ++p) {
            const float a = orc_h2f(data[p]);
            const float *b = Bf + (size_t)idx[p] * N;
            for (int j = 0; j < N; ++j) c[j] += a * b[j];
        }
    }
}

void orc_csr_spmm_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int M, const float *Bf, int N,
                      float *C) {
    memset(C, 0, sizeof(float) * (size_t)M * (size_t)N);
    orc_csr_spmm_acc_f32(hdr, idx, data, M, Bf, N, C);
}

/* ---------------------------------------------------------------------------
 * FP8 E4M3 quantisation (storage-mode extension; no reference counterpart).
 * ------------------------------------------------------------------------ */
uint8_t orc_f2e4m3(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint8_t sign = (uint8_t)((x >> 24) & 0x80u);
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return sign | 0x7fu;  /* NaN */
    if (ax >= 0x43e00000u) return sign | 0x7eu; /* >= 448 (and inf): saturate */
    if (ax < 0x3c800000u) {                     /* below 2^-6: multiples of 2^-9 */
        uint32_t e = ax >> 23;
        if (e < 116u) return sign;              /* below 2^-11: rounds to 0 */
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 141u - e;              /* value / 2^-9 = m * 2^(e - 141) */
        uint32_t r = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (r & 1u))) r++;
        return sign | (uint8_t)r;               /* r == 8 is the smallest normal */
    }
    uint32_t e = ax >> 23;
    uint32_t r = ((e - 120u) << 3) | ((ax >> 20) & 7u);
    uint32_t rem = ax & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) r++;
    if (r > 0x7eu) r = 0x7eu;
    return sign | (uint8_t)r;
}

float orc_e4m32f(uint8_t b) {
    int e = (b >> 3) & 15, m = b & 7;
    float v;
    if (e == 15 && m == 7) return NAN;
    if (e == 0)
        v = ldexpf((float)m, -9);
    else
        v = ldexpf(1.0f + (float)m / 8.0f, e - 7);
    return (b & 0x80) ? -v : v;
}

uint8_t orc_e8m0_for_absmax(float absmax) {
    if (!(absmax > 0.0f)) return 127;
    int k = -127;
    while (k < 127 && ldexpf(448.0f, k) < absmax) k++;
    return (uint8_t)(k + 127);
}

void orc_quant_e4m3_groups(const float *x, int64_t n, int64_t len, uint8_t *q, uint8_t *scale) {
    for (int64_t g = 0; g < n; ++g) {
        float amax = 0.0f;
        for (int64_t i = 0; i < len; ++i) amax = fmaxf(amax, fabsf(x[g * len + i]));
        uint8_t se = orc_e8m0_for_absmax(amax);
        scale[g] = se;
        for (int64_t i = 0; i < len; ++i) q[g * len + i] = orc_f2e4m3(ldexpf(x[g * len + i], 127 - (int)se));
    }
}

void orc_quant_e4m3_B(const float *B, int K, int N, uint8_t *q, uint8_t *scale) {
    const int slabs = K / 16, chunks = (N + 255) / 256, stride = (slabs + 15) & ~15;
    for (int h = 0; h < chunks; ++h) {
        int n0 = h * 256, n1 = n0 + 256 < N ? n0 + 256 : N;
        for (int s = 0; s < slabs; ++s) {
            float amax = 0.0f;
            for (int k = 16 * s; k < 16 * s + 16; ++k)
                for (int n = n0; n < n1; ++n) amax = fmaxf(amax, fabsf(B[(size_t)k * N + n]));
            uint8_t se = orc_e8m0_for_absmax(amax);
            scale[(size_t)h * stride + s] = se;
            for (int k = 16 * s; k < 16 * s + 16; ++k)
                for (int n = n0; n < n1; ++n)
                    q[(size_t)k * N + n] = orc_f2e4m3(ldexpf(B[(size_t)k * N + n], 127 - (int)se));
        }
    }
}

void orc_quant_e4m3_tiles(const int32_t *hdr, const int32_t *idx, const float *data, int block_rows, int bh,
                          int bw, float *out) {
    (void)idx;
    const int pieces = bw / 16;
    const size_t blk = (size_t)bh * bw;
    for (int R = 0; R < block_rows; ++R) {
        const int64_t k0 = hdr[R], nblk = (int64_t)(hdr[R + 1] - hdr[R]) * pieces;
        for (int t = 0; t < bh / 16; ++t)
            for (int64_t d = 0; d < nblk; d += 2) {
                /* device blocks d and d + 1 (16 rows x 16 columns each) share a scale */
                float amax = 0.0f;
                for (int64_t e = d; e < d + 2 && e < nblk; ++e) {
                    const int64_t k = k0 + e / pieces;
                    const int q = (int)(e % pieces);
                    for (int i = 0; i < 16; ++i)
                        for (int c = 0; c < 16; ++c)
                            amax = fmaxf(amax, fabsf(data[k * blk + (size_t)(16 * t + i) * bw + 16 * q + c]));
                }
                const int se = orc_e8m0_for_absmax(amax);
                for (int64_t e = d; e < d + 2 && e < nblk; ++e) {
                    const int64_t k = k0 + e / pieces;
                    const int q = (int)(e % pieces);
                    for (int i = 0; i < 16; ++i)
                        for (int c = 0; c < 16; ++c) {
                            const size_t o = k * blk + (size_t)(16 * t + i) * bw + 16 * q + c;
                            out[o] = ldexpf(orc_e4m32f(orc_f2e4m3(ldexpf(data[o], 127 - se))), se - 127);
                        }
                }
            }
    }
}

/* One FP32 scale per block (s = absmax / 448, rounded to fp32), values
 * quantised as E4M3(x / s) and dequantised as E4M3 * s in fp32. */
void orc_quant_e4m3_blocks_f32scale(const float *data, int64_t nblk, int64_t blk, float *out) {
    for (int64_t b = 0; b < nblk; ++b) {
        const float *x = data + b * blk;
        float *o = out + b * blk;
        float amax = 0.0f;
        for (int64_t i = 0; i < blk; ++i) amax = fmaxf(amax, fabsf(x[i]));
        if (!(amax > 0.0f)) {
            for (int64_t i = 0; i < blk; ++i) o[i] = 0.0f;
            continue;
        }
        const float s = amax / 448.0f;
        for (int64_t i = 0; i < blk; ++i) o[i] = orc_e4m32f(orc_f2e4m3(x[i] / s)) * s;
    }
}

/* bfloat16 and hi+lo fp16 storage (extensions). */
uint16_t orc_f2bf16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    uint32_t lsb = (x >> 16) & 1u;
    return (uint16_t)((x + 0x7fffu + lsb) >> 16);
}

float orc_bf162f(uint16_t b) {
    uint32_t x = (uint32_t)b << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void orc_hilo_split(float a, uint16_t *hi, uint16_t *lo) {
    *hi = orc_f2h(a);
    *lo = orc_f2h((a - orc_h2f(*hi)) * 2048.0f);
}

/* The same conversions over arrays. */
void orc_f2bf16_n(const float *x, int64_t n, uint16_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_f2bf16(x[i]);
}

void orc_hilo_split_n(const float *x, int64_t n, uint16_t *hi, uint16_t *lo) {
    for (int64_t i = 0; i < n; ++i) orc_hilo_split(x[i], hi + i, lo + i);
}
