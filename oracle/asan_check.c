/* asan_check.c -- runs the CPU oracle's entry points on small inputs in a
 * build with the address and undefined-behaviour sanitizers (test
 * infrastructure, like the oracle itself). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int orc_arpc(const double* X, int64_t n, int32_t dim, double eps, int64_t rank, int32_t b, uint64_t seed,
             double tol, int32_t max_rounds, const int64_t* inject, int64_t n_inject, double* F, int64_t* pivots,
             int32_t* accepted, double* err, int32_t* rounds_out, int64_t* rank_out, double* d_out,
             double* phase_s);
void orc_bandwidth(const double* X, int64_t n, int32_t dim, const int64_t* rows, int64_t nrows,
                   const double* eps, int32_t neps, double* S);
int orc_propose(const double* d, int64_t n, uint64_t seed, uint32_t round, int32_t b, int64_t* out,
                uint64_t* W_out);

int main(void)
{
    const int64_t n = 300, rank = 60;
    const int32_t dim = 5, b = 16, mr = 64;
    double* X = malloc(sizeof(double) * n * dim);
    srand(3);
    for (int64_t i = 0; i < n * dim; ++i) X[i] = (double)rand() / RAND_MAX - 0.5;
    double* F = calloc((size_t)(n * rank), sizeof(double));
    int64_t* piv = calloc((size_t)rank, sizeof(int64_t));
    int32_t* acc = calloc((size_t)mr, sizeof(int32_t));
    double* err = calloc((size_t)mr, sizeof(double));
    double* d = calloc((size_t)n, sizeof(double));
    double ph[3];
    int32_t rounds = 0;
    int64_t rk = 0;
    int rc = orc_arpc(X, n, dim, 0.3, rank, b, 1, 1e-10, mr, NULL, 0, F, piv, acc, err, &rounds, &rk, d, ph);
    int bad = rc != 0 || rk != rank;
    /* the same pivots injected reproduce the factor */
    double* F2 = calloc((size_t)(n * rank), sizeof(double));
    rc = orc_arpc(X, n, dim, 0.3, rank, b, 1, 1e-10, mr, piv, rank, F2, piv, acc, err, &rounds, &rk, d, NULL);
    for (int64_t i = 0; i < n * rank; ++i) bad |= fabs(F[i] - F2[i]) > 1e-9;
    int64_t rows[7] = {0, 5, 17, 42, 99, 150, 299};
    double eps[3] = {0.1, 1.0, 10.0}, S[3];
    orc_bandwidth(X, n, dim, rows, 7, eps, 3, S);
    bad |= !(S[0] >= 7.0 && S[2] <= 7.0 * n);
    int64_t out[8];
    uint64_t W = 0;
    bad |= orc_propose(d, n, 9, 0, 8, out, &W) != 0 && W != 0;
    for (int j = 0; j < 8; ++j) bad |= W != 0 && (out[j] < 0 || out[j] >= n);
    printf("oracle asan_check: %s\n", bad ? "FAILED" : "ok");
    free(X); free(F); free(F2); free(piv); free(acc); free(err); free(d);
    return bad;
}
