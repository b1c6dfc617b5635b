/*
 * C restatement of the reference's dense-statevector gate semantics.
 *
 * TEST INFRASTRUCTURE ONLY: a checker and the CPU baseline of the benchmark,
 * never part of the product.  It follows oracle/sv_oracle.py (which cites the
 * reference file:line of every rule) and is pinned against it by
 * tests/test_oracle_c.py.  Written independently of the product's C++ and
 * HIP sources.
 *
 * Conventions (reference backends/statevector.py:22-33): the state is 2^n
 * interleaved (re, im) numbers, qubit q is bit n-1-q of the basis index.  A
 * k-qubit matrix acting on qubits[0..k) is indexed with qubits[0] as its most
 * significant bit (statevector.py:188-203 for k = 2: index 2*bit(q1)+bit(q2)).
 * Op-table rows are [g_id, q1, q2, param] (circuits/compiler.py:24-48).
 *
 * Work is split over OpenMP threads by groups of 2^k amplitudes.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cd;

/* 2x2 matrix of 1-qubit gate ids 0..14 (sv_oracle.gate_1q_matrix). */
static int gate_1q(int g, double p, cd m[4]) {
    const double r = 1.0 / sqrt(2.0), c = cos(p / 2.0), s = sin(p / 2.0);
    const cd e = cexp(I * p);
    cd t[4] = {1, 0, 0, 1};
    switch (g) {
        case 0: break;
        case 1: t[0] = r; t[1] = r; t[2] = r; t[3] = -r; break;
        case 2: t[0] = 0; t[1] = 1; t[2] = 1; t[3] = 0; break;
        case 3: t[0] = 0; t[1] = -I; t[2] = I; t[3] = 0; break;
        case 4: t[3] = -1; break;
        case 5: t[3] = I; break;
        case 6: t[3] = -I; break;
        case 7: t[3] = cexp(I * M_PI / 4.0); break;
        case 8: t[3] = cexp(-I * M_PI / 4.0); break;
        case 9: t[0] = c; t[1] = -I * s; t[2] = -I * s; t[3] = c; break;
        case 10: t[0] = c; t[1] = -s; t[2] = s; t[3] = c; break;
        case 11: t[0] = cexp(-0.5 * I * p); t[3] = cexp(0.5 * I * p); break;
        case 12: t[3] = e; break;
        case 13: t[0] = 0.5 + 0.5 * I; t[1] = 0.5 - 0.5 * I; t[2] = 0.5 - 0.5 * I; t[3] = 0.5 + 0.5 * I; break;
        case 14: t[0] = e; t[3] = e; break;
        default: return -1;
    }
    memcpy(m, t, sizeof(t));
    return 0;
}

/* 4x4 matrix (index 2*bit(q1)+bit(q2)) of 2-qubit ids 20..25
 * (sv_oracle.gate_2q_matrix). */
static int gate_2q(int g, double p, cd m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1 : 0;
    switch (g) {
        case 20: m[10] = 0; m[11] = 1; m[14] = 1; m[15] = 0; break;
        case 21: m[15] = -1; break;
        case 22: m[15] = cexp(I * p); break;
        case 23: m[5] = 0; m[6] = 1; m[9] = 1; m[10] = 0; break;
        case 24: m[10] = 0; m[11] = -I; m[14] = I; m[15] = 0; break;
        case 25: m[10] = cexp(-0.5 * I * p); m[15] = cexp(0.5 * I * p); break;
        default: return -1;
    }
    return 0;
}

/* Deposit the bits of x around the k bit positions in sorted[] (ascending):
 * the index of group x with zeros at those positions. */
static inline uint64_t spread_bits(uint64_t x, const int *sorted, int k) {
    for (int j = 0; j < k; ++j) {
        const int p = sorted[j];
        x = ((x >> p) << (p + 1)) | (x & ((1ull << p) - 1));
    }
    return x;
}

#define DEFINE_APPLY(NAME, REAL)                                                                   \
    static int NAME(REAL *sv, int n, const cd *mat, const int *qubits, int k, int threads) {      \
        if (k < 1 || k > 6 || n < k) return -1;                                                  \
        const int D = 1 << k;                                                                      \
        int bits[6], sorted[6];                                                                    \
        for (int j = 0; j < k; ++j) {                                                              \
            if (qubits[j] < 0 || qubits[j] >= n) return -1;                                        \
            bits[j] = n - 1 - qubits[j];                                                           \
            sorted[j] = bits[j];                                                                   \
        }                                                                                          \
        for (int a = 0; a < k; ++a)                                                                \
            for (int b = a + 1; b < k; ++b) {                                                      \
                if (sorted[b] == sorted[a]) return -1;                                             \
                if (sorted[b] < sorted[a]) { int t = sorted[a]; sorted[a] = sorted[b]; sorted[b] = t; } \
            }                                                                                      \
        uint64_t off[64];                                                                          \
        for (int r = 0; r < D; ++r) {                                                              \
            uint64_t o = 0;                                                                        \
            for (int j = 0; j < k; ++j)                                                            \
                if (r >> (k - 1 - j) & 1) o |= 1ull << bits[j];                                    \
            off[r] = o;                                                                            \
        }                                                                                          \
        const int64_t groups = (int64_t)1 << (n - k);                                              \
        _Pragma("omp parallel for num_threads(threads) schedule(static)")                          \
        for (int64_t g = 0; g < groups; ++g) {                                                     \
            const uint64_t base = spread_bits((uint64_t)g, sorted, k);                             \
            cd a[64];                                                                              \
            for (int r = 0; r < D; ++r) {                                                          \
                const REAL *x = sv + 2 * (base | off[r]);                                          \
                a[r] = (double)x[0] + I * (double)x[1];                                            \
            }                                                                                      \
            for (int r = 0; r < D; ++r) {                                                          \
                cd y = 0;                                                                          \
                for (int c = 0; c < D; ++c) y += mat[r * D + c] * a[c];                            \
                REAL *x = sv + 2 * (base | off[r]);                                                \
                x[0] = (REAL)creal(y);                                                             \
                x[1] = (REAL)cimag(y);                                                             \
            }                                                                                      \
        }                                                                                          \
        return 0;                                                                                  \
    }

DEFINE_APPLY(apply_f64, double)
DEFINE_APPLY(apply_f32, float)

/* sv (complex128 or complex64 per `single`) <- U sv; mat is row-major
 * 2^k x 2^k complex128 (interleaved doubles).  Returns 0 or -1. */
int svo_apply_matrix(void *sv, int single, int n, const double *mat, const int *qubits, int k, int threads) {
    const cd *m = (const cd *)mat;
    return single ? apply_f32((float *)sv, n, m, qubits, k, threads)
                  : apply_f64((double *)sv, n, m, qubits, k, threads);
}

/* Apply an op table [g_id, q1, q2, param] row by row. */
int svo_run_ops(void *sv, int single, int n, const double *ops, int64_t n_ops, int threads) {
    for (int64_t i = 0; i < n_ops; ++i) {
        const int g = (int)ops[4 * i];
        const int q1 = (int)ops[4 * i + 1], q2 = (int)ops[4 * i + 2];
        const double p = ops[4 * i + 3];
        cd m[16];
        int rc;
        if (g < 20) {
            if (gate_1q(g, p, m)) return -1;
            const int q[1] = {q1};
            rc = svo_apply_matrix(sv, single, n, (const double *)m, q, 1, threads);
        } else {
            if (gate_2q(g, p, m)) return -1;
            const int q[2] = {q1, q2};
            rc = svo_apply_matrix(sv, single, n, (const double *)m, q, 2, threads);
        }
        if (rc) return rc;
    }
    return 0;
}

/* Threads OpenMP would use by default (1 without OpenMP). */
int svo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
