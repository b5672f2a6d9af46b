// ksgen.cpp -- Kuramoto-Sivashinsky snapshot generator.
//
// u_t = -u u_x - u_xx - u_xxxx on [0, L) periodic, M grid points, ETDRK4
// (Kassam & Trefethen 2005) in Fourier space with 3/2-rule dealiasing and
// 16-point contour-integral coefficients: the scheme of ks_datagen.py:5-89.
// The initial condition is ic_scale * N(0,1) per grid point (Box-Muller on
// Philox4x32-10, counter {p/2, 0, GAUSS, 0}), n_discard steps are dropped and
// one snapshot is kept every `stride` steps.
//
// The real transforms are applied as dense matrices over the M/2+1 retained
// modes: the padded inverse transform of length Mext = 3M/2 and the forward
// transform truncated to the retained modes.  With M <= 256 this is a few
// thousand fused multiply-adds per transform on contiguous rows, which the
// compiler vectorises; the trajectory is inherently sequential in time.
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "kevd.h"

namespace {

typedef std::complex<double> cplx;

void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
            uint32_t out[4])
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct KS {
    int M, M2, Mext;
    std::vector<double> e1, e2, f0, f1, f2, f3;
    std::vector<double> ggi;        // gg = i * ggi (purely imaginary)
    // padded inverse: u[x] = sum_j Binv[j][x] * c_j over the 2*M2 real
    // coefficients c = (Re v_0, Im v_0, Re v_1, Im v_1, ...)
    std::vector<double> Binv;       // (2*M2) x Mext
    // truncated forward: W[j] = sum_x Bfwd[x][j] w[x], j over 2*M2 reals
    std::vector<double> Bfwd;       // Mext x RP (RP = 2*M2 padded to 16, zero columns)
    int RP;
    std::vector<double> Bout;       // (2*M2) x M, the M-point inverse transform
    std::vector<double> u, w, W;
    std::vector<double> Nu, Na, Nb, Nc, a, b, c;

    KS(double L, int M_, double dt) : M(M_), M2(M_ / 2 + 1), Mext(3 * M_ / 2)
    {
        const double PI = 3.14159265358979323846;
        e1.resize(M2); e2.resize(M2); f0.resize(M2); f1.resize(M2); f2.resize(M2); f3.resize(M2);
        ggi.resize(M2);
        const int J = 16;
        for (int k = 0; k < M2; ++k) {
            double q = 2.0 * PI / L * k;
            double ld = q * q - q * q * q * q;
            ggi[k] = -0.5 * (double)Mext / (double)M * q;
            e1[k] = std::exp(dt * ld);
            e2[k] = std::exp(dt / 2 * ld);
            cplx s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int j = 0; j < J; ++j) {
                double th = PI * (0.5 + j) / J;
                cplx z = dt * ld + cplx(std::cos(th), std::sin(th));
                cplx ez = std::exp(z), ez2 = std::exp(0.5 * z), z3 = z * z * z;
                s0 += (ez2 - 1.0) / z;
                s1 += (-4.0 - z + ez * (4.0 - 3.0 * z + z * z)) / z3;
                s2 += 2.0 * (2.0 + z + ez * (-2.0 + z)) / z3;
                s3 += (-4.0 - 3.0 * z - z * z + ez * (4.0 - z)) / z3;
            }
            f0[k] = dt * std::real(s0) / J;
            f1[k] = dt * std::real(s1) / J;
            f2[k] = dt * std::real(s2) / J;
            f3[k] = dt * std::real(s3) / J;
        }
        int R = 2 * M2;
        RP = (R + 15) / 16 * 16;
        Binv.assign((size_t)R * Mext, 0.0);
        Bfwd.assign((size_t)Mext * RP, 0.0);
        for (int x = 0; x < Mext; ++x) {
            for (int k = 0; k < M2; ++k) {
                double a = 2.0 * PI * (double)k * (double)x / (double)Mext;
                double c = std::cos(a), s = std::sin(a);
                // irfft(v, Mext): (1/Mext)(Re v0 + 2 sum_{k>=1} Re(v_k e^{i a}));
                // Im v0 is ignored; all retained k < Mext/2.
                double wk = (k == 0 ? 1.0 : 2.0) / (double)Mext;
                Binv[(size_t)(2 * k) * Mext + x] = wk * c;
                Binv[(size_t)(2 * k + 1) * Mext + x] = k == 0 ? 0.0 : -wk * s;
                // rfft(w)[k] = sum_x w[x] e^{-i a}
                Bfwd[(size_t)x * RP + 2 * k] = c;
                Bfwd[(size_t)x * RP + 2 * k + 1] = -s;
            }
        }
        // irfft(v, M): (1/M)(Re v0 + 2 sum_{0<k<M/2} Re(v_k e^{i a}) + Re v_{M/2} cos(pi x))
        Bout.assign((size_t)R * M, 0.0);
        for (int x = 0; x < M; ++x)
            for (int k = 0; k < M2; ++k) {
                double ang = 2.0 * PI * (double)k * (double)x / (double)M;
                if (k == 0) {
                    Bout[(size_t)0 * M + x] = 1.0 / M;
                } else if (k == M / 2) {
                    Bout[(size_t)(2 * k) * M + x] = ((x & 1) ? -1.0 : 1.0) / M;
                } else {
                    Bout[(size_t)(2 * k) * M + x] = 2.0 * std::cos(ang) / M;
                    Bout[(size_t)(2 * k + 1) * M + x] = -2.0 * std::sin(ang) / M;
                }
            }
        u.resize(Mext);
        w.resize(RP);
        W.resize(RP);
        Nu.resize(R); Na.resize(R); Nb.resize(R); Nc.resize(R);
        a.resize(R); b.resize(R); c.resize(R);
    }

    // out = gg .* rfft(irfft(v, Mext)^2)[:M2]; v, out as interleaved re/im.
    // Both products run in blocks of 16 outputs held in registers (Mext and
    // RP are multiples of 16), so each matrix row is streamed once.
    void nonlin(const double* __restrict v, double* __restrict out)
    {
        const int R = 2 * M2;
        double* __restrict uu = u.data();
        const double* __restrict Bi = Binv.data();
        for (int xb = 0; xb < Mext; xb += 16) {
            double acc[16] = {};
            for (int j = 0; j < R; ++j) {
                double cj = v[j];
                const double* row = Bi + (size_t)j * Mext + xb;
#pragma GCC unroll 16
                for (int t = 0; t < 16; ++t) acc[t] += cj * row[t];
            }
#pragma GCC unroll 16
            for (int t = 0; t < 16; ++t) uu[xb + t] = acc[t] * acc[t];
        }
        double* __restrict WW = W.data();
        const double* __restrict Bf = Bfwd.data();
        for (int jb = 0; jb < RP; jb += 16) {
            double acc[16] = {};
            for (int x = 0; x < Mext; ++x) {
                double q = uu[x];
                const double* row = Bf + (size_t)x * RP + jb;
#pragma GCC unroll 16
                for (int t = 0; t < 16; ++t) acc[t] += q * row[t];
            }
#pragma GCC unroll 16
            for (int t = 0; t < 16; ++t) WW[jb + t] = acc[t];
        }
        // multiply by i*ggi: (a + ib) * i g = -b g + i a g
        for (int k = 0; k < M2; ++k) {
            double a = WW[2 * k], b = WW[2 * k + 1];
            out[2 * k] = -b * ggi[k];
            out[2 * k + 1] = a * ggi[k];
        }
    }

    void step(std::vector<double>& v)
    {
        nonlin(v.data(), Nu.data());
        for (int k = 0; k < M2; ++k)
            for (int q = 0; q < 2; ++q) a[2 * k + q] = e2[k] * v[2 * k + q] + f0[k] * Nu[2 * k + q];
        nonlin(a.data(), Na.data());
        for (int k = 0; k < M2; ++k)
            for (int q = 0; q < 2; ++q) b[2 * k + q] = e2[k] * v[2 * k + q] + f0[k] * Na[2 * k + q];
        nonlin(b.data(), Nb.data());
        for (int k = 0; k < M2; ++k)
            for (int q = 0; q < 2; ++q)
                c[2 * k + q] = e2[k] * a[2 * k + q] + f0[k] * (2.0 * Nb[2 * k + q] - Nu[2 * k + q]);
        nonlin(c.data(), Nc.data());
        for (int k = 0; k < M2; ++k)
            for (int q = 0; q < 2; ++q) {
                int j = 2 * k + q;
                v[j] = e1[k] * v[j] + f1[k] * Nu[j] + f2[k] * (Na[j] + Nb[j]) + f3[k] * Nc[j];
            }
    }

    void to_physical(const std::vector<double>& v, double* out) const
    {
        const int R = 2 * M2;
        for (int x = 0; x < M; ++x) out[x] = 0.0;
        for (int j = 0; j < R; ++j) {
            double cj = v[j];
            const double* row = &Bout[(size_t)j * M];
            for (int x = 0; x < M; ++x) out[x] += cj * row[x];
        }
    }

    // rfft of a real M-vector
    void to_spectral(const double* u0, std::vector<double>& v) const
    {
        const double PI = 3.14159265358979323846;
        v.assign(2 * M2, 0.0);
        for (int k = 0; k < M2; ++k) {
            double re = 0.0, im = 0.0;
            for (int x = 0; x < M; ++x) {
                double a = 2.0 * PI * (double)k * (double)x / (double)M;
                re += u0[x] * std::cos(a);
                im -= u0[x] * std::sin(a);
            }
            v[2 * k] = re;
            v[2 * k + 1] = im;
        }
    }
};

}  // namespace

extern "C" kevd_status kevd_ks_generate(const kevd_ks_params* p, int64_t n, double* out)
{
    if (!p || !out || n < 1 || p->M < 4 || p->M > 1024 || (p->M % 32) != 0 || !(p->L > 0.0) ||
        !(p->dt > 0.0) || p->n_discard < 0 || p->stride < 1)
        return KEVD_EINVAL;
    KS ks(p->L, p->M, p->dt);
    std::vector<double> u0(p->M);
    for (int q = 0; q < p->M; q += 2) {
        uint32_t o[4];
        philox((uint32_t)(q / 2), 0u, 2u, 0u, (uint32_t)p->seed, (uint32_t)(p->seed >> 32), o);
        uint64_t a = ((((uint64_t)o[1]) << 32) | o[0]) >> 11;
        uint64_t b = ((((uint64_t)o[3]) << 32) | o[2]) >> 11;
        double u1 = (double)(a + 1) * 0x1.0p-53, u2 = (double)b * 0x1.0p-53;
        double rad = std::sqrt(-2.0 * std::log(u1));
        const double PI = 3.14159265358979323846;
        u0[q] = p->ic_scale * rad * std::cos(2.0 * PI * u2);
        if (q + 1 < p->M) u0[q + 1] = p->ic_scale * rad * std::sin(2.0 * PI * u2);
    }
    std::vector<double> v;
    ks.to_spectral(u0.data(), v);
    for (int64_t s = 0; s < p->n_discard; ++s) ks.step(v);
    ks.to_physical(v, out);
    for (int64_t i = 1; i < n; ++i) {
        for (int s = 0; s < p->stride; ++s) ks.step(v);
        ks.to_physical(v, out + i * p->M);
    }
    return KEVD_OK;
}

extern "C" kevd_status kevd_delay_embed(const double* series, int64_t ndata, int32_t M, int32_t nq,
                                        double* out)
{
    if (!series || !out || M < 1 || nq < 1 || ndata < nq) return KEVD_EINVAL;
    int64_t N = ndata - nq + 1;
    for (int32_t g = 0; g < M; ++g)
        for (int64_t t = 0; t < N; ++t) {
            double* o = out + ((int64_t)g * N + t) * nq;
            for (int32_t j = 0; j < nq; ++j) o[j] = series[(t + j) * M + g];
        }
    return KEVD_OK;
}
