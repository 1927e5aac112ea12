// Checker for the product's lammps cell from cell vectors (test
// infrastructure, built only where the reference's sources are present, into
// oracle/_ref; never shipped or timed).
//
// The reference turns a cell given as three row vectors into a LAMMPS cell
// and a rotation with Eigen (triclinic.h:10-64): the HouseholderQR of the
// cell matrix, Q from householderQ(), R's upper triangle transposed, then
// the columns of R^T and Q with a negative diagonal entry multiplied by
// diag(..,-1,..); positions and velocities are row vectors times Q.  This
// program runs those Eigen calls with the Eigen 3.3.9 headers the reference
// vendors (eigen3/), compiled as the reference's Release build compiles them
// (-O3, no -march).
//
// stdin: int32 count, then per frame 9 doubles (the numpy row-major cell)
// and int32 nvec and nvec*3 doubles; stdout per frame: the 9 LAMMPS box
// values (0, 0, 0, lx/2, ly/2, lz/2, xy, xz, yz), Q's 9 entries in
// column-major order and the nvec rotated vectors.
#include <Eigen/Dense>
#include <cstdint>
#include <cstdio>
#include <vector>

static bool read_all(void *p, size_t bytes) { return std::fread(p, 1, bytes, stdin) == bytes; }

int main() {
    int32_t count = 0;
    if (!read_all(&count, sizeof count) || count < 0) return 1;
    std::vector<double> out;
    for (int32_t f = 0; f < count; f++) {
        double buf[9];
        int32_t nvec = 0;
        if (!read_all(buf, sizeof buf) || !read_all(&nvec, sizeof nvec) || nvec < 0) return 1;
        std::vector<double> vec((size_t)nvec * 3);
        if (!read_all(vec.data(), vec.size() * sizeof(double))) return 1;
        Eigen::Map<Eigen::Matrix<double, 3, 3>> cell(buf);
        Eigen::Matrix<double, 3, 3> q, r;
        if (cell(0, 1) == 0 && cell(0, 2) == 0 && cell(1, 0) == 0 && cell(1, 2) == 0 && cell(2, 0) == 0 &&
            cell(2, 1) == 0) {
            q = Eigen::Matrix<double, 3, 3>::Identity();
            r = cell;
        } else {
            auto qr = cell.transpose().householderQr();
            q = qr.householderQ();
            r = qr.matrixQR().triangularView<Eigen::Upper>().transpose();
            for (int i = 0; i < 3; ++i)
                if (r(i, i) < 0) {
                    Eigen::Matrix<double, 3, 3> s = Eigen::Matrix<double, 3, 3>::Identity();
                    s(i, i) = -1;
                    r = r * s;
                    q = q * s;
                }
        }
        const double box[9] = {0, 0, 0, r(0, 0) / 2, r(1, 1) / 2, r(2, 2) / 2, r(1, 0), r(2, 0), r(2, 1)};
        out.insert(out.end(), box, box + 9);
        out.insert(out.end(), q.data(), q.data() + 9);
        for (int32_t n = 0; n < nvec; n++) {
            Eigen::Map<Eigen::Matrix<double, 1, 3>> v(&vec[(size_t)n * 3]);
            v = v * q;
            out.insert(out.end(), &vec[(size_t)n * 3], &vec[(size_t)n * 3] + 3);
        }
    }
    return std::fwrite(out.data(), sizeof(double), out.size(), stdout) == out.size() ? 0 : 1;
}
