// su3gf_config.cpp -- host-side generator of synthetic gauge configurations
// (the benchmark inputs; not a reference interface).
//
//   U_mu(x) = exp(i sum_a c_a lambda_a / 2),  c_a ~ N(0, sigma^2)
//
// with the Gell-Mann matrices lambda_a, followed (haar_seed != 0) by a
// Haar-random gauge transformation U_mu(x) <- g(x) U_mu(x) g(x+mu)^dagger.
// Random numbers: Philox4x32-10 keyed by (link or site number, seed), turned
// into normals with Box-Muller.  The output is the reference's GpuPattern
// layout (18 reals per link, parity-split site index), FP64, on the host.
// Deterministic for a given (dims, sigma, seed, haar_seed) whatever the
// thread count.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/su3gf.h"

namespace {

typedef std::complex<double> cd;
struct Mat {
    cd a[3][3];
};

Mat mul(const Mat &x, const Mat &y)
{
    Mat r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cd s = 0;
            for (int k = 0; k < 3; k++) s += x.a[i][k] * y.a[k][j];
            r.a[i][j] = s;
        }
    return r;
}

Mat dag(const Mat &x)
{
    Mat r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.a[i][j] = std::conj(x.a[j][i]);
    return r;
}

void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// a stream of uniforms in (0,1) keyed by (id, seed) within domain `dom`
struct Stream {
    uint32_t k0, k1, dom, n = 0;
    double buf[2];
    int left = 0;
    double uniform()
    {
        if (left == 0) {
            uint32_t c[4] = {n++, dom, 0x13198A2Eu, 0x03707344u};
            philox10(c, k0, k1);
            const uint64_t w0 = (uint64_t)c[0] | ((uint64_t)c[1] << 32), w1 = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
            buf[0] = (double)(1 + (w0 >> 11)) * 0x1.fffffffffffffp-54;
            buf[1] = (double)(1 + (w1 >> 11)) * 0x1.fffffffffffffp-54;
            left = 2;
        }
        return buf[--left];
    }
    void normal2(double &z0, double &z1)
    {
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        z0 = r * std::cos(2.0 * M_PI * u2);
        z1 = r * std::sin(2.0 * M_PI * u2);
    }
};

// exp(i H) for Hermitian traceless H: scaling and squaring with a degree-16
// Taylor polynomial, then a Gram-Schmidt reunitarization (det = 1 by the
// cross-product third row).
Mat exp_iH(const Mat &H)
{
    double nrm = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) nrm += std::norm(H.a[i][j]);
    nrm = std::sqrt(nrm);
    int sq = 0;
    while (nrm > 0.25) {
        nrm *= 0.5;
        sq++;
    }
    const double scale = std::ldexp(1.0, -sq);
    Mat A;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A.a[i][j] = cd(0, 1) * H.a[i][j] * scale;
    Mat R;  // Horner: I + A(I + A/2(I + A/3(...)))
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R.a[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 16; k >= 1; k--) {
        Mat T = mul(A, R);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R.a[i][j] = ((i == j) ? 1.0 : 0.0) + T.a[i][j] / (double)k;
    }
    for (int s = 0; s < sq; s++) R = mul(R, R);
    return R;
}

void gram_schmidt_su3(Mat &u)
{
    double n0 = 0;
    for (int k = 0; k < 3; k++) n0 += std::norm(u.a[0][k]);
    n0 = 1.0 / std::sqrt(n0);
    for (int k = 0; k < 3; k++) u.a[0][k] *= n0;
    cd sp = 0;
    for (int k = 0; k < 3; k++) sp += u.a[1][k] * std::conj(u.a[0][k]);
    for (int k = 0; k < 3; k++) u.a[1][k] -= u.a[0][k] * sp;
    double n1 = 0;
    for (int k = 0; k < 3; k++) n1 += std::norm(u.a[1][k]);
    n1 = 1.0 / std::sqrt(n1);
    for (int k = 0; k < 3; k++) u.a[1][k] *= n1;
    u.a[2][0] = std::conj(u.a[0][1] * u.a[1][2] - u.a[0][2] * u.a[1][1]);
    u.a[2][1] = std::conj(u.a[0][2] * u.a[1][0] - u.a[0][0] * u.a[1][2]);
    u.a[2][2] = std::conj(u.a[0][0] * u.a[1][1] - u.a[0][1] * u.a[1][0]);
}

Mat random_link(uint32_t link, uint32_t seed, double sigma)
{
    Stream st{link, seed, 0x243F6A88u};
    double c[8];
    for (int a = 0; a < 8; a += 2) st.normal2(c[a], c[a + 1]);
    for (int a = 0; a < 8; a++) c[a] *= sigma;
    const double r3 = 1.0 / std::sqrt(3.0);
    Mat H;
    H.a[0][0] = 0.5 * (c[2] + c[7] * r3);
    H.a[1][1] = 0.5 * (-c[2] + c[7] * r3);
    H.a[2][2] = -c[7] * r3;
    H.a[0][1] = 0.5 * cd(c[0], -c[1]);
    H.a[0][2] = 0.5 * cd(c[3], -c[4]);
    H.a[1][2] = 0.5 * cd(c[5], -c[6]);
    H.a[1][0] = std::conj(H.a[0][1]);
    H.a[2][0] = std::conj(H.a[0][2]);
    H.a[2][1] = std::conj(H.a[1][2]);
    Mat U = exp_iH(H);
    gram_schmidt_su3(U);
    return U;
}

// Haar SU(3): Gram-Schmidt of two complex Gaussian rows, third row = conj(r0 x r1)
Mat haar_su3(uint32_t site, uint32_t seed)
{
    Stream st{site, seed, 0x85A308D3u};
    Mat g;
    for (int i = 0; i < 2; i++)
        for (int k = 0; k < 3; k++) {
            double x, y;
            st.normal2(x, y);
            g.a[i][k] = cd(x, y);
        }
    gram_schmidt_su3(g);
    return g;
}

int64_t parity_index(const int dims[4], const int c[4], int64_t V)
{
    const int64_t n = (((int64_t)c[0] * dims[1] + c[1]) * dims[2] + c[2]) * dims[3] + c[3];
    return ((c[0] + c[1] + c[2] + c[3]) % 2 == 0) ? n / 2 : n / 2 + V / 2;
}

}  // namespace

extern "C" gf_status gf_make_config(const int dims[4], double sigma, uint32_t seed, uint32_t haar_seed, double *U)
{
    if (!dims || !U) return GF_ERR_BAD_ARG;
    int64_t V = 1;
    for (int i = 0; i < 4; i++) {
        if (dims[i] <= 0 || (dims[i] % 2 && !(i == 0 && dims[i] == 1))) return GF_ERR_BAD_SHAPE;
        V *= dims[i];
    }
    if (!(sigma >= 0)) return GF_ERR_BAD_ARG;
    std::vector<Mat> g;
    if (haar_seed) {
        g.resize((size_t)V);
#pragma omp parallel for schedule(static)
        for (int64_t n = 0; n < V; n++) g[(size_t)n] = haar_su3((uint32_t)n, haar_seed);
    }
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < V; n++) {
        int c[4];
        int64_t r = n;
        for (int d = 3; d >= 0; d--) {
            c[d] = (int)(r % dims[d]);
            r /= dims[d];
        }
        const int64_t idx = parity_index(dims, c, V);
        for (int mu = 0; mu < 4; mu++) {
            Mat u;
            if (sigma > 0) u = random_link((uint32_t)(4 * n + mu), seed, sigma);
            else
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) u.a[i][j] = (i == j) ? 1.0 : 0.0;
            if (haar_seed) {
                int cf[4] = {c[0], c[1], c[2], c[3]};
                cf[mu] = (cf[mu] + 1) % dims[mu];
                const int64_t nf = (((int64_t)cf[0] * dims[1] + cf[1]) * dims[2] + cf[2]) * dims[3] + cf[3];
                u = mul(mul(g[(size_t)n], u), dag(g[(size_t)nf]));
                gram_schmidt_su3(u);
            }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    U[idx + V * (0 + 2 * (j + 3 * (i + 3 * mu)))] = u.a[i][j].real();
                    U[idx + V * (1 + 2 * (j + 3 * (i + 3 * mu)))] = u.a[i][j].imag();
                }
        }
    }
    return GF_OK;
}
