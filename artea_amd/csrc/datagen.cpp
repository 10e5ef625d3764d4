// Seeded synthetic workloads (host side).  Every value is a pure function of
// (seed, stream, index), so any thread count, any machine with the same
// libm, and any [first, first+n) window produce the same bytes.
//
//   kind 0: 256-component Gaussian mixture in D dims, centres U[0,128]^D,
//           isotropic sigma 12, uniform component; nearbyint, clamp [0,255].
//   kind 1: zero-mean Gaussian with eigenvalues i^-0.8 (i = 1..D) on a random
//           orthogonal basis (modified Gram-Schmidt QR of a seeded Gaussian
//           matrix, FP64), unit-normalised.
//   kind 2: 1024-component mixture, centres N(0,1)^D, sigma 0.3,
//           unit-normalised.
// Streams: 0 = centres / basis, 1 = base component ids, 2 = base noise,
// 3 = query component ids and noise.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "host.hpp"

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t idx) {
    return mix64(mix64(seed ^ mix64(stream + 0x51ull)) + idx);
}
inline double uniform01(uint64_t x) { return double((x >> 11) + 1) * 0x1.0p-53; }  // (0, 1]
// Box-Muller, cosine branch; two draws per normal.
inline double normal(uint64_t seed, uint64_t stream, uint64_t idx) {
    const double u1 = uniform01(draw(seed, stream, 2 * idx));
    const double u2 = uniform01(draw(seed, stream, 2 * idx + 1));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

std::vector<double> centres(int kind, uint64_t seed, uint32_t comps, uint32_t dim) {
    std::vector<double> c(size_t(comps) * dim);
    for (size_t i = 0; i < c.size(); ++i)
        c[i] = kind == 0 ? 128.0 * uniform01(draw(seed, 0, i)) : normal(seed, 0, i);
    return c;
}

std::vector<double> basis(uint64_t seed, uint32_t dim) {  // columns orthonormal, row-major dim x dim
    std::vector<double> a(size_t(dim) * dim);
    for (size_t i = 0; i < a.size(); ++i) a[i] = normal(seed, 0, i);
    for (uint32_t j = 0; j < dim; ++j) {  // modified Gram-Schmidt over columns
        for (uint32_t k = 0; k < j; ++k) {
            double dot = 0;
            for (uint32_t i = 0; i < dim; ++i) dot += a[size_t(i) * dim + k] * a[size_t(i) * dim + j];
            for (uint32_t i = 0; i < dim; ++i) a[size_t(i) * dim + j] -= dot * a[size_t(i) * dim + k];
        }
        double nrm = 0;
        for (uint32_t i = 0; i < dim; ++i) nrm += a[size_t(i) * dim + j] * a[size_t(i) * dim + j];
        nrm = std::sqrt(nrm);
        for (uint32_t i = 0; i < dim; ++i) a[size_t(i) * dim + j] /= nrm;
    }
    return a;
}

}  // namespace

extern "C" int artea_generate(int kind, uint64_t seed, uint32_t which, uint64_t first, uint64_t n, uint32_t dim,
                              uint32_t dim_pad, float* out, int nthreads) {
    ARTEA_CHECK_ARG(kind >= 0 && kind <= 2, "artea_generate: kind must be 0, 1 or 2");
    ARTEA_CHECK_ARG(dim > 0 && dim <= dim_pad && dim_pad % 16 == 0, "artea_generate: need 0 < dim <= dim_pad, dim_pad % 16 == 0");
    ARTEA_CHECK_ARG(out != nullptr || n == 0, "artea_generate: null output");
    const uint64_t noise_stream = which == 0 ? 2 : 3;
    const uint64_t comp_stream = which == 0 ? 1 : 3;
    const uint32_t comps = kind == 0 ? 256 : 1024;
    std::vector<double> c, q;
    std::vector<double> sq;
    if (kind == 1) {
        q = basis(seed, dim);
        sq.resize(dim);
        for (uint32_t i = 0; i < dim; ++i) sq[i] = std::sqrt(std::pow(double(i + 1), -0.8));
    } else {
        c = centres(kind, seed, comps, dim);
    }
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        std::vector<double> x(dim), z(dim);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < int64_t(n); ++r) {
            const uint64_t row = first + uint64_t(r);
            float* o = out + uint64_t(r) * dim_pad;
            // Query rows use a disjoint index range of the shared stream 3:
            // component id at 2^62 + row, noise from row * dim.
            const uint64_t noise_base = row * dim;
            if (kind == 0 || kind == 2) {
                const uint64_t cid_draw = which == 0 ? draw(seed, comp_stream, row)
                                                     : draw(seed, comp_stream, (1ull << 62) + row);
                const uint32_t comp = uint32_t(((cid_draw >> 32) * comps) >> 32);
                const double* cc = c.data() + size_t(comp) * dim;
                const double sigma = kind == 0 ? 12.0 : 0.3;
                for (uint32_t d = 0; d < dim; ++d) x[d] = cc[d] + sigma * normal(seed, noise_stream, noise_base + d);
                if (kind == 0) {
                    for (uint32_t d = 0; d < dim; ++d) {
                        double v = std::nearbyint(x[d]);
                        o[d] = float(v < 0 ? 0 : (v > 255 ? 255 : v));
                    }
                } else {
                    double nrm = 0;
                    for (uint32_t d = 0; d < dim; ++d) nrm += x[d] * x[d];
                    nrm = std::sqrt(nrm);
                    for (uint32_t d = 0; d < dim; ++d) o[d] = float(x[d] / nrm);
                }
            } else {
                for (uint32_t d = 0; d < dim; ++d) z[d] = sq[d] * normal(seed, noise_stream, noise_base + d);
                double nrm = 0;
                for (uint32_t i = 0; i < dim; ++i) {
                    double s = 0;
                    for (uint32_t k = 0; k < dim; ++k) s += q[size_t(i) * dim + k] * z[k];
                    x[i] = s;
                    nrm += s * s;
                }
                nrm = std::sqrt(nrm);
                for (uint32_t d = 0; d < dim; ++d) o[d] = float(x[d] / nrm);
            }
            for (uint32_t d = dim; d < dim_pad; ++d) o[d] = 0.0f;
        }
    }
    return ARTEA_OK;
}
