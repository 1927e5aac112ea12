// Checker for the oracle's schur00 (test infrastructure, built only where the
// reference's sources are present, into oracle/_ref; never shipped or timed).
//
// The reference's Green-Kubo Schur complement of the (0,0) entry
// (greenkuboNcomponentionicfluid.cpp:309-316, and :338-345 for the Einstein
// form) is one Eigen expression over a column-major map of the symmetrised
// integral matrix: c00 - c(0,1:) * inverse(c(1:,1:)) * c(1:,0).  Its bits are
// those of Eigen's evaluation of that expression (a dynamic-size inverse by
// partial-pivoting LU, then the two products), so this program evaluates the
// same expression with the Eigen 3.3.9 headers the reference vendors
// (eigen3/), compiled as the reference's Release build compiles them (-O3, no
// -march: SSE2 doubles, no FMA).
//
// stdin: int32 count, then per matrix int32 n and n*n doubles (column-major);
// stdout: count doubles, the (0,0) Schur complement of each.
#include <Eigen/Dense>
#include <cstdint>
#include <cstdio>
#include <vector>

static bool read_all(void *p, size_t bytes) { return std::fread(p, 1, bytes, stdin) == bytes; }

int main() {
    int32_t count = 0;
    if (!read_all(&count, sizeof count) || count < 0) return 1;
    std::vector<double> out;
    out.reserve((size_t)count);
    for (int32_t m = 0; m < count; m++) {
        int32_t n = 0;
        if (!read_all(&n, sizeof n) || n < 1) return 1;
        std::vector<double> buf((size_t)n * n);
        if (!read_all(buf.data(), buf.size() * sizeof(double))) return 1;
        Eigen::Map<Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic>> c(buf.data(), n, n);
        double k;
        if (n > 1)
            k = (c.block(0, 0, 1, 1) - c.block(0, 1, 1, n - 1) * c.block(1, 1, n - 1, n - 1).inverse() *
                                           c.block(1, 0, n - 1, 1))(0, 0);
        else
            k = c(0, 0);
        out.push_back(k);
    }
    return std::fwrite(out.data(), sizeof(double), out.size(), stdout) == out.size() ? 0 : 1;
}
