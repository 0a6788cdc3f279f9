// Hilbert order of cells (sfc.h).
#include "sfc.h"

#include <algorithm>
#include <numeric>

namespace mhdg
{
namespace
{
constexpr int kKeyBits = 20; // bits per axis, as the reference's partitioner
}

uint64_t hilbert_key(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t X[3] = {x, y, z};
    const uint32_t top = 1u << (kKeyBits - 1);
    // inverse undo of the Gray-code rotations, from the top bit down
    for (uint32_t q = top; q > 1; q >>= 1)
    {
        const uint32_t low = q - 1;
        for (int i = 0; i < 3; ++i)
        {
            if (X[i] & q)
                X[0] ^= low; // invert the low bits of the first axis
            else
            {
                const uint32_t swap = (X[0] ^ X[i]) & low; // exchange the low bits of the first axis and axis i
                X[0] ^= swap;
                X[i] ^= swap;
            }
        }
    }
    // Gray encode
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t flip = 0;
    for (uint32_t q = top; q > 1; q >>= 1)
        if (X[2] & q) flip ^= q - 1;
    for (uint32_t &v : X) v ^= flip;
    // transpose -> index: one bit of every axis per level, x first
    uint64_t key = 0;
    for (int b = kKeyBits - 1; b >= 0; --b)
        key = (key << 3) | (uint64_t((X[0] >> b) & 1u) << 2) | (uint64_t((X[1] >> b) & 1u) << 1) | uint64_t((X[2] >> b) & 1u);
    return key;
}

std::vector<int32_t> hilbert_order(const double *c, int32_t n)
{
    std::vector<int32_t> order(size_t(n > 0 ? n : 0));
    if (n <= 0) return order;
    double lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
    for (int32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], c[3 * size_t(i) + k]);
            hi[k] = std::max(hi[k], c[3 * size_t(i) + k]);
        }
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    const double scale = double((1u << kKeyBits) - 1) / ext;
    std::vector<uint64_t> key(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i)
        key[size_t(i)] = hilbert_key(static_cast<uint32_t>((c[3 * size_t(i)] - lo[0]) * scale), static_cast<uint32_t>((c[3 * size_t(i) + 1] - lo[1]) * scale),
                                     static_cast<uint32_t>((c[3 * size_t(i) + 2] - lo[2]) * scale));
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[size_t(a)] != key[size_t(b)] ? key[size_t(a)] < key[size_t(b)] : a < b; });
    return order;
}
} // namespace mhdg
