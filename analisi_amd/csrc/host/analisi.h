// C++ operator API of the time-lagged pair-correlation path, kept as in
// rikigigi/analisi v0.5.2 (same class names, constructor arguments, reset /
// calculate / calc_single_th / calc_end, buffer layouts and error messages) so
// that code written against the reference compiles against this header; the
// work is done on the GPU through the C ABI in analisi_hip.h.
//
//   VectorOp<T,TFLOAT>                operazionisulista.h:25-148
//   Trajectory_numpy                  trajectory_numpy.{h,cpp} (plain arrays in
//                                     place of pybind11 buffers)
//   lammps_bin::dump                  basetrajectory.cpp:4-47 (LAMMPS-binary writer;
//                                     the reader, Trajectory, is lammps_traj.h)
//   ReadLog_numpy<TFLOAT>             readlog_numpy.{h,cpp}
//   Gofrt<TFLOAT,T>                   gofrt.{h,cpp}
//   MSD<T,FPE>                        msd.{h,cpp}
//   SphericalCorrelations<T>          sphericalcorrelations.{h,cpp}, sphericalbase.{h,cpp}
//   Steinhardt<T>                     steinhardt.{h,cpp}
//   GreenKuboNComponentIonicFluid<..> greenkuboNcomponentionicfluid.{h,cpp}
//   MediaVar<T>, MediaVarCovar<T>     calcoliblocchi.h
//   BlockAverageG<TR,T,Args...>       blockaverage.h:83-218
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "analisi_hip.h"

namespace analisi {

inline void anl_check(int rc) {
    if (rc != ANL_OK) throw std::runtime_error(anl_last_error());
}

// ---------------------------------------------------------------------------
// VectorOp (operazionisulista.h:25-148): owner of vdata, element-wise ops
// ---------------------------------------------------------------------------
template <class T, class TFLOAT = double>
class VectorOp {
   public:
    VectorOp &operator=(const VectorOp &destra) {
        if (data_length != destra.data_length) {
            delete[] vdata;
            data_length = destra.data_length;
            vdata = new TFLOAT[data_length];
        }
        for (unsigned int i = 0; i < data_length; i++) vdata[i] = destra.vdata[i];
        return *this;
    }
#define ANL_VOP(OP)                                                                         \
    T &operator OP(const T &destra) {                                                       \
        if (destra.lunghezza() != data_length)                                              \
            throw std::runtime_error("Trying to operate on VectorOp of different sizes!");  \
        for (unsigned int i = 0; i < data_length; i++) vdata[i] OP destra.elemento(i);      \
        return static_cast<T &>(*this);                                                     \
    }                                                                                       \
    T &operator OP(const TFLOAT &destra) {                                                  \
        for (unsigned int i = 0; i < data_length; i++) vdata[i] OP destra;                  \
        return static_cast<T &>(*this);                                                     \
    }
    ANL_VOP(+=)
    ANL_VOP(-=)
    ANL_VOP(*=)
    ANL_VOP(/=)
#undef ANL_VOP
    unsigned int lunghezza() const { return data_length; }
    TFLOAT elemento(unsigned int i) const {
        if (i < data_length) return vdata[i];
        throw std::runtime_error("Out of range index");
    }
    TFLOAT *access_vdata() { return vdata; }
    void azzera() {
        for (unsigned int i = 0; i < data_length; i++) vdata[i] = 0;
    }

   protected:
    VectorOp() : vdata(nullptr), data_length(0) {}
    VectorOp(const VectorOp &c) : vdata(nullptr), data_length(0) { operator=(c); }
    ~VectorOp() { delete[] vdata; }
    TFLOAT *vdata;
    unsigned int data_length;
};

// ---------------------------------------------------------------------------
// LAMMPS-binary writer (basetrajectory.cpp:4-47), shared by both trajectories
// ---------------------------------------------------------------------------
namespace lammps_bin {

constexpr int NDOUBLE_ATOMO = 8;  // id, type, x, y, z, vx, vy, vz (lammps_struct.h:104-122)

// internal box (xlo, ylo, zlo, lx/2, ly/2, lz/2) -> lammps (basetrajectory.h:110-119)
inline void internal_to_lammps(double *c) {
    double t = c[3];
    c[3] = c[4] * 2 + c[1];
    c[5] = c[5] * 2 + c[2];
    c[4] = c[2];
    c[2] = c[1];
    c[1] = c[0] + t * 2;
}

// Timesteps [start_ts, stop_ts) (all of them when stop_ts <= 0), one chunk
// each, the old header layout (lammps_struct.h:65-97), atom ids 0..N-1 and
// the internal type index as the type.
template <class T>
void dump(T &t, const std::string &fname, int start_ts, int stop_ts) {
    const int64_t nts = (int64_t)t.get_ntimesteps();
    if (start_ts < 0 || start_ts >= nts)
        throw std::runtime_error("You must provide a starting timestep between 0 and the number of timesteps!");
    if (stop_ts <= 0) stop_ts = (int)nts;
    std::ofstream out(fname, std::ofstream::binary);
    const int64_t natoms = (int64_t)t.get_natoms();
    const int tri = t.is_triclinic() ? 1 : 0;
    for (int ts = start_ts; ts < stop_ts; ++ts) {
        double box[6], tilt[3] = {0, 0, 0};
        const double *b = t.box(ts);
        for (int i = 0; i < 6; ++i) box[i] = b[i];
        if (tri)
            for (int i = 0; i < 3; ++i) tilt[i] = b[6 + i];
        internal_to_lammps(box);
        const int64_t step = ts;
        const int zero6[6] = {0, 0, 0, 0, 0, 0};
        const int size_one = NDOUBLE_ATOMO, nchunk = 1;
        out.write((const char *)&step, sizeof(step));
        out.write((const char *)&natoms, sizeof(natoms));
        out.write((const char *)&tri, sizeof(tri));
        out.write((const char *)zero6, sizeof(zero6));
        out.write((const char *)box, sizeof(box));
        if (tri) out.write((const char *)tilt, sizeof(tilt));
        out.write((const char *)&size_one, sizeof(size_one));
        out.write((const char *)&nchunk, sizeof(nchunk));
        const int n_data = (int)(natoms * NDOUBLE_ATOMO);
        out.write((const char *)&n_data, sizeof(n_data));
        for (int64_t ia = 0; ia < natoms; ++ia) {
            double data[NDOUBLE_ATOMO];
            data[0] = (double)ia;
            data[1] = (double)t.get_type((unsigned int)ia);
            const double *x = t.positions(ts, (int)ia);
            for (int k = 0; k < 3; ++k) data[2 + k] = x[k];
            const double *v = t.velocity(ts, (int)ia);
            for (int k = 0; k < 3; ++k) data[5 + k] = v[k];
            out.write((const char *)data, sizeof(data));
        }
    }
    if (!out) throw std::runtime_error("Error writing the trajectory \"" + fname + "\"");
}

}  // namespace lammps_bin

// ---------------------------------------------------------------------------
// Trajectory_numpy (trajectory_numpy.cpp:7-199)
// ---------------------------------------------------------------------------
class Trajectory_numpy {
   public:
    enum class BoxFormat { Lammps_ortho, Lammps_triclinic, Cell_vectors, Invalid };

    // pos/vel [nsteps][natoms][3], types [natoms], box [nsteps][6|9] (lammps)
    // or [nsteps][3][3] (cell vectors, row-major as numpy gives it).
    Trajectory_numpy(const double *pos, const double *vel, const int *types, const double *box, size_t nsteps,
                     size_t natoms, BoxFormat fmt, bool wrap_pbc = false, bool save_rotation_matrix = false)
        : n_timesteps(nsteps), natoms(natoms), wrap_pbc(wrap_pbc), box_format(fmt) {
        if (fmt == BoxFormat::Invalid) throw std::runtime_error("Invalid input cell format");
        triclinic = false;
        buffer_boxes_stride = 6;
        std::vector<double> rot;
        if (fmt == BoxFormat::Lammps_ortho || fmt == BoxFormat::Lammps_triclinic) {
            if (fmt == BoxFormat::Lammps_triclinic) {
                buffer_boxes_stride = 9;
                triclinic = true;
            }
            boxes.assign(box, box + nsteps * buffer_boxes_stride);
            for (size_t i = 0; i < nsteps; ++i) lammps_to_internal(&boxes[i * buffer_boxes_stride]);
        } else {
            // triclinic.h:10-64: lammps cell from the QR of the cell matrix
            std::vector<Lcell> cells(nsteps);
            for (size_t i = 0; i < nsteps; i++) {
                cells[i] = lammps_cell(box + 9 * i);
                if (!cells[i].diagonal) triclinic = true;
            }
            buffer_boxes_stride = triclinic ? 9 : 6;
            boxes.assign(nsteps * buffer_boxes_stride, 0.0);
            for (size_t i = 0; i < nsteps; i++) {
                double *c = &boxes[i * buffer_boxes_stride];
                c[3] = cells[i].r[0][0] / 2;
                c[4] = cells[i].r[1][1] / 2;
                c[5] = cells[i].r[2][2] / 2;
                if (triclinic) {
                    c[6] = cells[i].r[1][0];
                    c[7] = cells[i].r[2][0];
                    c[8] = cells[i].r[2][1];
                }
            }
            if (triclinic) {
                rot.resize(nsteps * 9);
                for (size_t i = 0; i < nsteps; i++) std::memcpy(&rot[9 * i], cells[i].q, sizeof(double) * 9);
                rotations = cells;
            }
        }
        pos_.assign(pos, pos + nsteps * natoms * 3);
        vel_.assign(vel, vel + nsteps * natoms * 3);
        if (triclinic && fmt == BoxFormat::Cell_vectors) {
            for (size_t i = 0; i < nsteps; i++)
                for (size_t n = 0; n < natoms; n++) {
                    rotate(rotations[i], &pos_[(i * natoms + n) * 3]);
                    rotate(rotations[i], &vel_[(i * natoms + n) * 3]);
                }
            if (save_rotation_matrix) rotation_matrix = rot;
        }
        raw_types.assign(types, types + natoms);
        compute_type_ids();
        calc_cm(pos, positions_cm_);
        calc_cm(vel, velocities_cm_);
        if (wrap_pbc)
            for (size_t i = 0; i < nsteps; ++i) pbc_wrap(i);
    }
    ~Trajectory_numpy() {
        if (dev) anl_traj_destroy(dev);
    }
    Trajectory_numpy(const Trajectory_numpy &) = delete;
    Trajectory_numpy &operator=(const Trajectory_numpy &) = delete;

    double *positions(int t, int i) { return &pos_[(size_t)t * natoms * 3 + (size_t)i * 3]; }
    double *velocity(int t, int i) { return &vel_[(size_t)t * natoms * 3 + (size_t)i * 3]; }
    double *box(int t) { return &boxes[(size_t)t * buffer_boxes_stride]; }
    double *positions_cm(int t, int type) { return &positions_cm_[((size_t)t * ntypes + type) * 3]; }
    double *velocity_cm(int t, int type) { return &velocities_cm_[((size_t)t * ntypes + type) * 3]; }
    double *get_rotation_matrix(size_t t) { return rotation_matrix.empty() ? nullptr : &rotation_matrix[9 * t]; }
    unsigned int get_type(unsigned int i) const {
        if (i < natoms) return type_id[i];
        throw std::runtime_error("Atom index out of range\n");
    }
    std::vector<unsigned int> get_types() const { return types; }
    size_t get_ntypes() const { return ntypes; }
    size_t get_natoms() const { return natoms; }
    size_t get_ntimesteps() const { return n_timesteps; }
    size_t get_nloaded_timesteps() const { return n_timesteps; }
    ssize_t get_current_timestep() const { return 0; }
    bool get_pbc_wrap() const { return wrap_pbc; }
    bool is_triclinic() const { return triclinic; }
    size_t get_box_stride() const { return buffer_boxes_stride; }
    void dump_lammps_bin_traj(const std::string &fname, int start_ts, int stop_ts) {
        lammps_bin::dump(*this, fname, start_ts, stop_ts);
    }
    // every timestep is resident: a calculation on [primo, primo + need)
    // uses device timestep primo; the contents never change
    size_t window_offset(size_t primo, size_t /*need*/) { return primo; }
    uint64_t generation() const { return 0; }

    // basetrajectory.h:94-105
    static void lammps_to_internal(double *c) {
        double t = c[4];
        c[5] = (c[5] - c[4]) / 2;
        c[4] = (c[3] - c[2]) / 2;
        c[3] = (c[1] - c[0]) / 2;
        c[1] = c[2];
        c[2] = t;
    }
    // basetrajectory.h:224-268
    static void minImage_triclinic(double *d, const double *lh, const double *t, bool tri) {
        double xy = tri ? t[0] : 0, xz = tri ? t[1] : 0, yz = tri ? t[2] : 0;
        while (std::fabs(d[2]) > lh[2]) {
            if (d[2] < 0.0) { d[2] += lh[2] * 2; if (tri) { d[1] += yz; d[0] += xz; } }
            else { d[2] -= lh[2] * 2; if (tri) { d[1] -= yz; d[0] -= xz; } }
        }
        while (std::fabs(d[1]) > lh[1]) {
            if (d[1] < 0.0) { d[1] += lh[1] * 2; if (tri) d[0] += xy; }
            else { d[1] -= lh[1] * 2; if (tri) d[0] -= xy; }
        }
        while (std::fabs(d[0]) > lh[0]) {
            if (d[0] < 0.0) d[0] += lh[0] * 2;
            else d[0] -= lh[0] * 2;
        }
    }
    // basetrajectory.h:181-219, on the host (pyanalisi's Trajectory.minImage)
    double d2_minImage(size_t i, size_t j, size_t ti, size_t tj, double *x) {
        const double *c = box((int)ti);
        for (int k = 0; k < 3; k++) x[k] = positions((int)ti, (int)i)[k] - positions((int)tj, (int)j)[k];
        minImage_triclinic(x, c + 3, c + 6, triclinic);
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) d2 += x[k] * x[k];
        return d2;
    }

    // device mirror, built on first use and immutable afterwards
    anl_traj *device() {
        if (!dev) {
            std::vector<int32_t> tid(type_id.begin(), type_id.end());
            anl_check(anl_traj_create(pos_.data(), boxes.data(), (uint32_t)buffer_boxes_stride, tid.data(),
                                      (uint32_t)ntypes, positions_cm_.data(), n_timesteps, natoms, &dev));
        }
        return dev;
    }

   private:
    struct Lcell {
        double q[9];     // column-major Q
        double r[3][3];  // lower-triangular lammps cell (rows: a, b, c)
        bool diagonal;
    };
    std::vector<Lcell> rotations;
    // triclinic.h:10-40: cell is the row-major numpy [3][3] seen column-major;
    // the QR of its transpose (the numpy matrix itself, rows = cell vectors)
    // by Householder reflections, then the signs of R's diagonal made
    // positive.  The reference calls Eigen's HouseholderQR on a fixed 3 x 3
    // matrix, so the steps and their rounding follow that computation (Eigen
    // 3.3.9, the version the reference vendors): per column k the reflector
    // (makeHouseholder: beta = -sign(c0) |x|, essential = tail / (c0 - beta),
    // tau = (beta - c0) / beta, a zero tail leaving the column as it is),
    // applied to the columns on its right (tmp = essential . rows below +
    // row k; row k -= tau tmp; rows below -= tmp (tau essential)); Q from the
    // identity with the reflectors applied last to first; the sign flips as
    // products with diag(.., -1, ..).  Sums here have at most two terms
    // before the row term, so their order is Eigen's.
    static void apply_reflector(double (&M)[3][3], int k, const double (&A)[3][3], double tau, int c0, int c1) {
        // rows k..2 of columns c0..c1 of M, the reflector's essential part in A(k+1.., k)
        for (int j = c0; j <= c1; j++) {
            double t = 0;
            if (k == 0)
                t = A[1][0] * M[1][j] + A[2][0] * M[2][j];
            else if (k == 1)
                t = A[2][1] * M[2][j];
            t += M[k][j];
            M[k][j] -= tau * t;
            for (int i = k + 1; i < 3; i++) M[i][j] -= t * (tau * A[i][k]);
        }
    }
    // r = r diag(s) (s_i = -1 or 1) as a 3 x 3 product, sums in column order
    static void times_sign(double (&r)[3][3], int i) {
        double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        m[i][i] = -1;
        double o[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) o[a][b] = r[a][0] * m[0][b] + r[a][1] * m[1][b] + r[a][2] * m[2][b];
        std::memcpy(r, o, sizeof(o));
    }
    static Lcell lammps_cell(const double *m) {
        Lcell L{};
        double cell[3][3];  // cell(i,j) = m[j*3+i]
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cell[i][j] = m[j * 3 + i];
        if (cell[0][1] == 0 && cell[0][2] == 0 && cell[1][0] == 0 && cell[1][2] == 0 && cell[2][0] == 0 &&
            cell[2][1] == 0) {
            L.diagonal = true;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    L.r[i][j] = cell[i][j];
                    L.q[j * 3 + i] = i == j ? 1.0 : 0.0;
                }
            return L;
        }
        L.diagonal = false;
        double A[3][3], tau[3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = cell[j][i];
        for (int k = 0; k < 3; k++) {
            const double tail = k == 0 ? A[1][0] * A[1][0] + A[2][0] * A[2][0] : (k == 1 ? A[2][1] * A[2][1] : 0.0);
            const double c0 = A[k][k];
            double beta;
            if (tail <= std::numeric_limits<double>::min()) {
                tau[k] = 0;
                beta = c0;
                for (int i = k + 1; i < 3; i++) A[i][k] = 0;
            } else {
                beta = std::sqrt(c0 * c0 + tail);
                if (c0 >= 0) beta = -beta;
                for (int i = k + 1; i < 3; i++) A[i][k] = A[i][k] / (c0 - beta);
                tau[k] = (beta - c0) / beta;
            }
            A[k][k] = beta;
            if (k < 2 && tau[k] != 0) apply_reflector(A, k, A, tau[k], k + 1, 2);
        }
        double Q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int k = 1; k >= 0; k--)  // the last reflector is the identity (tau = 0)
            if (tau[k] != 0) apply_reflector(Q, k, A, tau[k], k, 2);
        double r[3][3];  // R's upper triangle, transposed (the lammps cell rows)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r[i][j] = j <= i ? A[j][i] : 0.0;
        for (int i = 0; i < 3; i++)  // triclinic.h:30-38 reflect
            if (r[i][i] < 0) {
                times_sign(r, i);
                times_sign(Q, i);
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                L.r[i][j] = r[i][j];
                L.q[j * 3 + i] = Q[i][j];
            }
        return L;
    }
    // triclinic.h:53-56: row vector times Q
    static void rotate(const Lcell &c, double *v) {
        double o[3];
        for (int j = 0; j < 3; j++) o[j] = v[0] * c.q[j * 3 + 0] + v[1] * c.q[j * 3 + 1] + v[2] * c.q[j * 3 + 2];
        std::memcpy(v, o, sizeof(o));
    }
    // basetrajectory.cpp:51-89
    void compute_type_ids() {
        std::vector<int> distinct;
        for (size_t i = 0; i < natoms; i++) {
            bool seen = false;
            for (int d : distinct) seen |= d == raw_types[i];
            if (!seen) distinct.push_back(raw_types[i]);
        }
        std::sort(distinct.begin(), distinct.end());
        std::map<int, unsigned int> m;
        for (unsigned int k = 0; k < distinct.size(); k++) {
            m[distinct[k]] = k;
            types.push_back(distinct[k]);
        }
        ntypes = distinct.size();
        type_id.resize(natoms);
        for (size_t i = 0; i < natoms; i++) type_id[i] = m.at(raw_types[i]);
    }
    // trajectory_numpy.cpp:201-225: running mean of the raw input per type
    void calc_cm(const double *a, std::vector<double> &cm) {
        cm.assign(n_timesteps * ntypes * 3, 0.0);
        std::vector<int> cont(ntypes);
        for (size_t i = 0; i < n_timesteps; ++i) {
            std::fill(cont.begin(), cont.end(), 0);
            for (size_t ia = 0; ia < natoms; ++ia) {
                unsigned int it = type_id[ia];
                cont[it]++;
                for (int k = 0; k < 3; k++)
                    cm[(i * ntypes + it) * 3 + k] +=
                        (a[(i * natoms + ia) * 3 + k] - cm[(i * ntypes + it) * 3 + k]) / double(cont[it]);
            }
        }
    }
    // basetrajectory.h:145-161
    void pbc_wrap(size_t idx) {
        double *c = box((int)idx);
        double x0[3] = {c[3], c[4], c[5]};
        for (size_t ia = 0; ia < natoms; ++ia) {
            double *xa = positions((int)idx, (int)ia);
            for (int k = 0; k < 3; ++k) xa[k] = xa[k] - x0[k];
            minImage_triclinic(xa, c + 3, c + 6, triclinic);
            for (int k = 0; k < 3; ++k) xa[k] = xa[k] + x0[k];
        }
    }

    size_t n_timesteps, natoms, ntypes = 0, buffer_boxes_stride;
    bool wrap_pbc, triclinic;
    BoxFormat box_format;
    std::vector<double> pos_, vel_, boxes, positions_cm_, velocities_cm_, rotation_matrix;
    std::vector<int> raw_types;
    std::vector<unsigned int> type_id, types;
    anl_traj *dev = nullptr;
};

// ---------------------------------------------------------------------------
// ReadLog_numpy (readlog_numpy.{h,cpp})
// ---------------------------------------------------------------------------
template <class TFLOAT = double>
class ReadLog_numpy {
   public:
    ReadLog_numpy(const TFLOAT *data, size_t nrows, size_t ncols, std::vector<std::string> headers)
        : headers(std::move(headers)), ntimestep((unsigned)nrows), ncols((unsigned)ncols), data_(data) {
        if (this->headers.size() != ncols)
            throw std::runtime_error("Number of headers is not equal to the number of columns");
    }
    ~ReadLog_numpy() {
        if (dev) anl_series_destroy(dev);
    }
    TFLOAT *line(unsigned int index) { return const_cast<TFLOAT *>(data_) + (size_t)index * ncols; }
    unsigned int n_timestep() { return ntimestep; }
    std::pair<unsigned int, bool> get_index_of(const std::string &header) {
        for (unsigned int i = 0; i < ncols; ++i)
            if (header == headers.at(i)) return {i, true};
        return {0, false};
    }
    anl_series *device() {
        if (!dev) anl_check(anl_series_create(data_, ntimestep, ncols, &dev));
        return dev;
    }

   private:
    std::vector<std::string> headers;
    unsigned int ntimestep, ncols;
    const TFLOAT *data_;
    anl_series *dev = nullptr;
};

// ---------------------------------------------------------------------------
// Gofrt (gofrt.h:28-107, gofrt.cpp)
// ---------------------------------------------------------------------------
// whether a trajectory holds frames a device handle can be built over now:
// a windowed trajectory only once its window has been read
template <class X, class = void>
struct windowed_reader : std::false_type {};
template <class X>
struct windowed_reader<X, std::void_t<decltype(std::declval<const X &>().has_window_data())>> : std::true_type {};
template <class TR>
inline bool has_frames(TR *t) {
    if constexpr (windowed_reader<TR>::value)
        return t->has_window_data();
    else
        return t->get_nloaded_timesteps() > 0;
}

template <class TFLOAT, class T>
class Gofrt : public VectorOp<Gofrt<TFLOAT, T>, TFLOAT> {
    using VOp = VectorOp<Gofrt<TFLOAT, T>, TFLOAT>;
    using VOp::data_length;
    using VOp::vdata;

   public:
    Gofrt(T *t, TFLOAT rmin, TFLOAT rmax, unsigned int nbin, unsigned int tmax = 0, unsigned int nthreads = 0,
          unsigned int skip = 1, unsigned int every = 1, bool debug = false, int form = ANL_GOFRT_AUTO)
        : traiettoria(t), rmin(rmin), rmax(rmax), nbin(nbin), lmax(tmax), nthreads(nthreads == 0 ? 1 : nthreads),
          skip(skip == 0 ? 1 : skip), every(every == 0 ? 1 : every), debug(debug), form(form) {
        if (has_frames(t)) h();  // argument errors surface here
    }
    ~Gofrt() {
        if (g) anl_gofrt_destroy(g);
    }
    Gofrt &operator=(const Gofrt &destra) {
        VOp::operator=(destra);
        return *this;
    }
    unsigned int nExtraTimesteps(unsigned int n_b) {  // gofrt.cpp:37-39
        unsigned int v = (unsigned int)(traiettoria->get_ntimesteps() / (n_b + 1) + 1);
        return (v < lmax || lmax == 0) ? v : lmax;
    }
    void reset(const unsigned int n) {  // gofrt.cpp:41-63
        c_descr = legend();
        leff = (n < lmax || lmax == 0) ? n : lmax;
        ntimesteps = n;
        data_length = leff * nch() * nbin;
        delete[] vdata;
        vdata = new TFLOAT[data_length];
        reset_done = true;
        if (g && gen == traiettoria->generation()) anl_check(anl_gofrt_reset(g, n, nullptr));
    }
    std::vector<ssize_t> get_shape() { return {(ssize_t)leff, (ssize_t)nch(), (ssize_t)nbin}; }
    std::vector<ssize_t> get_stride() {
        return {(ssize_t)(nch() * nbin * sizeof(TFLOAT)), (ssize_t)(nbin * sizeof(TFLOAT)), (ssize_t)sizeof(TFLOAT)};
    }
    std::string get_columns_description() { return c_descr; }
    // CalculateMultiThread::calculate(primo) with Gofrt's flags: calc_init,
    // the whole (lag, t0, pair) nest in one device call, calc_end.
    void calculate(size_t primo) {
        calc_init((int)primo);
        // timesteps primo .. primo + n - 1 + leff - 1 must be on the device
        const size_t off = traiettoria->window_offset(primo, (size_t)ntimesteps + leff - 1);
        std::vector<uint64_t> counts(data_length);
        anl_check(anl_gofrt_calculate(h(), off, counts.data()));
        for (unsigned int k = 0; k < data_length; k++) vdata[k] = (TFLOAT)counts[k] * incr;
        calc_end();
    }
    void calc_init(int primo) {  // gofrt.cpp:74-93
        if ((uint64_t)leff + ntimesteps + primo > traiettoria->get_ntimesteps() + 1)
            throw std::runtime_error(
                "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
                "the size of the average or the lenght of the time lag");
        VOp::azzera();
        incr = (ntimesteps / skip > 0) ? 1.0 / int(ntimesteps / skip) : 1;
    }
    // gofrt.cpp:96-122: one lag t and t0 index imedia, atoms [a0,a1); adds
    // incr per hit into vdata (the ith argument is kept for the signature).
    void calc_single_th(int t, int imedia, int a0, int a1, int primo, int /*ith*/) {
        std::vector<uint64_t> counts(data_length, 0);
        const size_t off = traiettoria->window_offset((size_t)primo, (size_t)imedia + (size_t)t + 1);
        anl_check(anl_gofrt_calc_single_th(h(), t, imedia, a0, a1, (uint64_t)off, counts.data()));
        for (unsigned int k = 0; k < data_length; k++)
            if (counts[k]) vdata[k] += (TFLOAT)counts[k] * incr;
    }
    void calc_end() {  // gofrt.cpp:127-155 (debug dump)
        if (!debug) return;
        std::ofstream out("gofrt.dump", std::ios::app);
        for (unsigned int ts = 0; ts < leff; ts++) {
            for (unsigned int r = 0; r < nbin; r++) {
                out << ts << " " << r;
                for (unsigned int it = 0; it < nch(); it++) out << " " << vdata[(size_t)ts * nch() * nbin + it * nbin + r];
                out << "\n";
            }
        }
        out << "\n\n";
    }
    anl_gofrt *handle() { return g; }

   private:
    // device handle over the trajectory's current device mirror; rebuilt
    // (with the same reset) when a windowed trajectory has moved its window
    anl_gofrt *h() {
        if (!g || gen != traiettoria->generation()) {
            anl_traj *d = traiettoria->device();
            if (g) anl_gofrt_destroy(g);
            g = nullptr;
            anl_check(anl_gofrt_create(d, rmin, rmax, nbin, lmax, skip, every, form, &g));
            gen = traiettoria->generation();
            if (reset_done) anl_check(anl_gofrt_reset(g, ntimesteps, nullptr));
        }
        return g;
    }
    unsigned int nch() const { return (unsigned int)(traiettoria->get_ntypes() * (traiettoria->get_ntypes() + 1)); }
    // gofrt.h:85-104: channel of the type pair (t1, t2), t1 <= t2, among the
    // nt (nt+1) / 2 distinct-atom channels; counted from the last pair back
    unsigned int pair_channel(unsigned int t1, unsigned int t2) const {
        const unsigned int nt = (unsigned int)traiettoria->get_ntypes();
        const unsigned int lo = std::min(t1, t2), hi = std::max(t1, t2);
        return nt * (nt + 1) / 2 - (hi + 1) * (hi + 2) / 2 + lo;
    }
    // gofrt.cpp:43-53: the command line's legend of its g(r,t) columns (1-based:
    // lag, bin, then a mean and a variance per channel)
    std::string legend() const {
        const unsigned int nt = (unsigned int)traiettoria->get_ntypes();
        std::ostringstream d;
        d << "# The first column is the time difference in timesteps, then you have the bin index. Every column "
             "after is followed by the variance. Then you have the following: \n";
        for (unsigned int a = 0; a < nt; a++)
            for (unsigned int b = a; b < nt; b++) {
                const unsigned int c = pair_channel(a, b);
                d << "#g(" << a << ", " << b << "), different atom index: " << c * 2 + 3 << "\n";
                d << "#g(" << a << ", " << b << "), same atom index: " << (c + nt * (nt + 1) / 2) * 2 + 3 << "\n";
            }
        d << "# same atom index means that the atom is tracked around and the average self-spread is shown with "
             "larger time differences.\n";
        d << "# different atom index is something that for t=0 is the traditional g(r) \n";
        return d.str();
    }
    T *traiettoria;
    TFLOAT rmin, rmax, incr = 1;
    unsigned int nbin, lmax, nthreads, skip, every;
    unsigned int leff = 0, ntimesteps = 0;
    bool debug;
    int form;
    bool reset_done = false;
    uint64_t gen = 0;
    anl_gofrt *g = nullptr;
    std::string c_descr;
};

// ---------------------------------------------------------------------------
// MSD (msd.h:26-60, msd.cpp)
// ---------------------------------------------------------------------------
template <class T, bool FPE = false>
class MSD : public VectorOp<MSD<T, FPE>> {
    using VOp = VectorOp<MSD<T, FPE>>;
    using VOp::data_length;
    using VOp::vdata;

   public:
    MSD(T *t, unsigned int skip = 1, unsigned int tmax = 0, unsigned int nthreads = 0, bool cm_msd = false,
        bool cm_self = false, bool debug = false)
        : traiettoria(t), lmax(tmax), f_cm(cm_msd ? 2 : 1), skip(skip == 0 ? 1 : skip), cm_msd(cm_msd),
          cm_self(cm_self), debug(debug) {}
    std::vector<ssize_t> get_shape() const { return {(ssize_t)leff, (ssize_t)f_cm, (ssize_t)ntypes}; }
    std::vector<ssize_t> get_stride() const {
        return {(ssize_t)(ntypes * f_cm * sizeof(double)), (ssize_t)(ntypes * sizeof(double)), (ssize_t)sizeof(double)};
    }
    void reset(const unsigned int n) {  // msd.cpp:40-49
        leff = (n < lmax || lmax == 0) ? n : lmax;
        ntypes = traiettoria->get_ntypes();
        data_length = leff * ntypes * f_cm;
        ntimesteps = n;
        delete[] vdata;
        vdata = new double[data_length];
    }
    void calculate(size_t primo) {
        calc_init(primo);
        const size_t off = traiettoria->window_offset(primo, leff + ntimesteps);
        anl_check(anl_msd_calculate(traiettoria->device(), skip, lmax, cm_msd, cm_self, ntimesteps, off, vdata));
        calc_end();
    }
    void calc_init(size_t primo) {  // msd.cpp:52-61
        if (leff + ntimesteps + primo > traiettoria->get_ntimesteps())
            throw std::runtime_error(
                "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
                "the size of the average or the lenght of the time lag");
        if (leff + ntimesteps > traiettoria->get_nloaded_timesteps()) {
            std::stringstream ss;
            ss << "there are not enough loaded timesteps inside the trajectory object. I need at least "
               << leff + ntimesteps << " timesteps to do the requested calculation";
            throw std::runtime_error(ss.str());
        }
    }
    void calc_end() {  // msd.cpp:129-146
        if (!debug) return;
        std::ofstream out("msd.dump", std::ios::app);
        for (size_t ts = 0; ts < leff; ts++) {
            out << ts;
            for (size_t it = 0; it < ntypes * f_cm; it++) out << " " << vdata[ntypes * ts * f_cm + it];
            out << "\n";
        }
        out << "\n\n";
    }
    MSD &operator=(const MSD &destra) {
        VOp::operator=(destra);
        return *this;
    }
    unsigned int nExtraTimesteps(unsigned int n_b) {  // msd.cpp:36-38
        unsigned int v = (unsigned int)(traiettoria->get_ntimesteps() / (n_b + 1) + 1);
        return (v < lmax || lmax == 0) ? v : lmax;
    }

   private:
    T *traiettoria;
    size_t lmax, f_cm, ntypes = 0, leff = 0, ntimesteps = 0;
    unsigned int skip;
    bool cm_msd, cm_self, debug;
};

// ---------------------------------------------------------------------------
// The device handle of SphericalBase (sphericalbase.h:6-60): ranges, bins and
// lmax over the trajectory's current device mirror, rebuilt when a windowed
// trajectory has moved its window.  The constructor checks the ranges as
// sphericalbase.cpp:7-16 does.
// ---------------------------------------------------------------------------
template <class T>
class SphericalHandle {
   public:
    using rminmax_t = std::vector<std::pair<double, double>>;
    using NeighListSpec = std::vector<std::tuple<size_t, double, double>>;  // neighbour.h:13
    SphericalHandle(T *t, const rminmax_t &rminmax, size_t nbin, unsigned int lmax, const NeighListSpec &nls = {})
        : traiettoria(t), nbin(nbin), lmax(lmax) {
        const size_t ntypes = t->get_ntypes();
        if (ntypes * ntypes != rminmax.size()) {
            std::stringstream ss;
            ss << "you must provide a radial range for each pair of atomic types, in total ntypes*ntypes pair of "
                  "numbers. You provided "
               << rminmax.size() << " elements while ntypes is " << ntypes << " .";
            throw std::runtime_error(ss.str());
        }
        for (const auto &p : rminmax) {
            rm.push_back(p.first);
            rm.push_back(p.second);
        }
        for (const auto &x : nls) {
            nneigh.push_back((uint32_t)std::get<0>(x));
            cut2.push_back(std::get<1>(x));
        }
    }
    ~SphericalHandle() {
        if (hs) anl_sph_destroy(hs);
    }
    SphericalHandle(const SphericalHandle &) = delete;
    SphericalHandle &operator=(const SphericalHandle &) = delete;
    anl_sph *get() {
        if (!hs || gen != traiettoria->generation()) {
            anl_traj *d = traiettoria->device();
            if (hs) anl_sph_destroy(hs);
            hs = nullptr;
            // neighbour.h:31-45: checked when the list is built, after the
            // calculation's own argument checks
            if (!nneigh.empty() && nneigh.size() != traiettoria->get_ntypes())
                throw std::runtime_error("In neighbours list you must specify parameters for each atomic type");
            anl_check(anl_sph_create(d, lmax, rm.data(), (uint32_t)nbin, &hs));
            if (!nneigh.empty()) anl_check(anl_sph_set_neighbours(hs, nneigh.data(), cut2.data()));
            gen = traiettoria->generation();
        }
        return hs;
    }

   private:
    T *traiettoria;
    std::vector<double> rm, cut2;
    std::vector<uint32_t> nneigh;
    size_t nbin;
    unsigned int lmax;
    anl_sph *hs = nullptr;
    uint64_t gen = 0;
};

// ---------------------------------------------------------------------------
// SphericalCorrelations (sphericalcorrelations.h:9-78, sphericalcorrelations.cpp)
// over SphericalBase (sphericalbase.h:6-60): the time correlation of the
// per-atom real spherical-harmonic densities of each neighbour type and
// distance bin.  vdata [leff][ntypes][ntypes][nbin][lmax+1], index k of the
// last axis holding l = lmax - k.  With a neighbour list (the SANN variant)
// the radial ranges are ignored and every atom's SANN neighbours of each type
// fill bin 0.
// ---------------------------------------------------------------------------
template <class T>
class SphericalCorrelations : public VectorOp<SphericalCorrelations<T>> {
    using VOp = VectorOp<SphericalCorrelations<T>>;
    using VOp::data_length;
    using VOp::vdata;

   public:
    using rminmax_t = std::vector<std::pair<double, double>>;
    using NeighListSpec = std::vector<std::tuple<size_t, double, double>>;  // neighbour.h:13
    SphericalCorrelations(T *t, const rminmax_t rminmax, size_t nbin, size_t tmax = 0, size_t nthreads = 0,
                          size_t skip = 1, size_t buffer_size = 10, bool debug = false,
                          const NeighListSpec neighList = {}, unsigned int lmax = ANL_SPH_LMAX_MAX)
        : traiettoria(t), handle(t, rminmax, nbin, lmax, neighList), nbin(nbin), tmax(tmax), nthreads(nthreads),
          skip(skip == 0 ? 1 : skip), buffer_size(buffer_size), lmax(lmax), ntypes(t->get_ntypes()) {
        (void)debug;
        if (has_frames(t)) handle.get();  // argument errors surface here
    }
    SphericalCorrelations &operator=(const SphericalCorrelations &destra) {
        VOp::operator=(destra);
        return *this;
    }
    unsigned int nExtraTimesteps(unsigned int n_b) {  // sphericalcorrelations.cpp:32-35
        size_t v = traiettoria->get_ntimesteps() / (n_b + 1) + 1;
        return (unsigned int)((v < tmax || tmax == 0) ? v : tmax);
    }
    void reset(const unsigned int n) {  // sphericalcorrelations.cpp:37-59
        ntimesteps = n;
        leff = (n < tmax || tmax == 0) ? n : tmax;
        data_length = leff * ntypes * ntypes * nbin * (lmax + 1);
        delete[] vdata;
        vdata = new double[data_length];
    }
    // sphericalcorrelations.cpp:62-236: the argument checks of the reference,
    // then the whole (t0, lag) nest in one device call
    void calculate(unsigned int primo) {
        if (leff + ntimesteps + primo > traiettoria->get_ntimesteps())
            throw std::runtime_error(
                "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
                "the size of the average or the lenght of the correlation function");
        if (nthreads <= 1) {
            std::cerr << "Warning: I'm using a single thread.\n";
            nthreads = 1;
        }
        if (buffer_size < 2) {
            std::cerr << "Warning: buffer size cannot be less than 2, setting it to 2" << std::endl;
            buffer_size = 2;
        }
        if (leff / nthreads == 0)
            throw std::runtime_error(
                "Don't have enough timesteps to parallelize the calculation. Try again with a smaller number of "
                "threads. ");
        const size_t off = traiettoria->window_offset(primo, ntimesteps + leff - 1);
        anl_check(anl_sph_calculate(handle.get(), off, (uint32_t)ntimesteps, (uint32_t)tmax, (uint32_t)skip, vdata));
    }
    std::vector<ssize_t> get_shape() const {
        return {(ssize_t)leff, (ssize_t)ntypes, (ssize_t)ntypes, (ssize_t)nbin, (ssize_t)(lmax + 1)};
    }
    std::vector<ssize_t> get_stride() const {  // sphericalcorrelations.h:40-49: C order
        auto s = get_shape();
        std::vector<ssize_t> st(s.size());
        ssize_t acc = sizeof(double);
        for (int i = (int)s.size() - 1; i >= 0; --i) {
            st[i] = acc;
            acc *= s[i];
        }
        return st;
    }
    std::string get_columns_description() { return "#TODO\n"; }
    // SphericalBase::calc(timestep) without counters (sphericalbase.cpp:19-66):
    // [natoms][ntypes][nbin][(lmax+1)^2], the reference's (l, m) order
    std::vector<double> calc(size_t timestep) {
        const size_t off = traiettoria->window_offset(timestep, 1);
        std::vector<double> out((size_t)traiettoria->get_natoms() * ntypes * nbin * (lmax + 1) * (lmax + 1));
        anl_check(anl_sph_snapshot(handle.get(), off, out.data()));
        return out;
    }
    size_t get_snap_size() const { return (size_t)traiettoria->get_natoms() * ntypes * nbin * (lmax + 1) * (lmax + 1); }

   private:
    T *traiettoria;
    SphericalHandle<T> handle;
    size_t nbin, tmax, nthreads, skip, buffer_size;
    unsigned int lmax;
    size_t ntypes, leff = 0, ntimesteps = 0;
};

// ---------------------------------------------------------------------------
// Steinhardt<l> (steinhardt.h:17-120, steinhardt.cpp:7-236): per frame, the
// order parameters q_l of every (atom, neighbour type, bin), as a histogram
// over the listed l values or as a trajectory of q_1..q_lmax.  With a neighbour
// list (SANN, one radial bin) the radial ranges are ignored, and the averaged
// order parameter is available.
// ---------------------------------------------------------------------------
template <class T>
class Steinhardt : public VectorOp<Steinhardt<T>> {
    using VOp = VectorOp<Steinhardt<T>>;
    using VOp::data_length;
    using VOp::vdata;

   public:
    using Rminmax_t = std::vector<std::pair<double, double>>;
    using NeighListSpec = std::vector<std::tuple<size_t, double, double>>;  // neighbour.h:13
    Steinhardt(T *t, const Rminmax_t rminmax, unsigned int nbin, unsigned int steinhardt_histogram_size,
               std::vector<unsigned int> steinhardt_l_histogram, unsigned int nthreads = 2, unsigned int skip = 1,
               bool do_histogram = true, const NeighListSpec nls = {}, const bool averaged_order = false,
               unsigned int lmax = ANL_SPH_LMAX_MAX)
        : traiettoria(t), handle(t, rminmax, nbin, lmax, nls), l_hist(std::move(steinhardt_l_histogram)), nbin(nbin),
          nbin_st(steinhardt_histogram_size), skip(skip == 0 ? 1 : skip), lmax(lmax), ntypes(t->get_ntypes()),
          do_histogram(do_histogram), averaged(averaged_order) {
        (void)nthreads;
        if (nls.empty() && averaged_order)  // steinhardt.cpp:30-35
            throw std::runtime_error(
                "You must use a neighbour list when computing a averaged steinhardt order parameter! ");
        if (!nls.empty() && nbin > 1)
            throw std::runtime_error("When using a neighbour list to compute steinhardt please set nbin to 1. ");
        hist_size = 1;
        for (size_t i = 0; i < l_hist.size(); ++i) hist_size *= nbin_st;
        if (has_frames(t)) handle.get();
    }
    Steinhardt &operator=(const Steinhardt &destra) {
        VOp::operator=(destra);
        return *this;
    }
    unsigned int nExtraTimesteps(unsigned int) { return 0; }
    void reset(const unsigned int n) {  // steinhardt.cpp:47-80
        ntimesteps = n;
        const size_t len = do_histogram ? hist_size * nbin * ntypes * ntypes
                                        : (size_t)nbin * (n / skip) * traiettoria->get_natoms() * ntypes * lmax;
        if (len != data_length) {
            delete[] vdata;
            data_length = (unsigned int)len;
            vdata = new double[data_length];
        }
    }
    void calculate(size_t primo) {  // calculatemultithread.h:106-162 with calc_init, steinhardt.cpp:80-94
        if (ntimesteps + primo > traiettoria->get_ntimesteps())
            throw std::runtime_error(
                "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
                "the size of the average ");
        const size_t off = traiettoria->window_offset(primo, ntimesteps);
        anl_check(anl_steinhardt_calculate(handle.get(), off, (uint32_t)ntimesteps, skip, l_hist.data(),
                                           (uint32_t)l_hist.size(), nbin_st, do_histogram ? 1 : 0, averaged ? 1 : 0,
                                           vdata));
    }
    std::vector<ssize_t> get_shape() const {  // steinhardt.h:76-87
        if (do_histogram) {
            std::vector<ssize_t> r{(ssize_t)nbin, (ssize_t)ntypes, (ssize_t)ntypes};
            for (size_t i = 0; i < l_hist.size(); ++i) r.push_back(nbin_st);
            return r;
        }
        return {(ssize_t)nbin, (ssize_t)(ntimesteps / skip), (ssize_t)traiettoria->get_natoms(), (ssize_t)ntypes,
                (ssize_t)lmax};
    }
    std::vector<ssize_t> get_stride() const {  // steinhardt.h:48-74: C order
        auto s = get_shape();
        std::vector<ssize_t> st(s.size());
        ssize_t acc = sizeof(double);
        for (int i = (int)s.size() - 1; i >= 0; --i) {
            st[i] = acc;
            acc *= s[i];
        }
        return st;
    }

   private:
    T *traiettoria;
    SphericalHandle<T> handle;
    std::vector<unsigned int> l_hist;
    unsigned int nbin, nbin_st, skip, lmax;
    size_t ntypes, hist_size = 1, ntimesteps = 0;
    bool do_histogram, averaged;
};

// ---------------------------------------------------------------------------
// GreenKuboNComponentIonicFluid (greenkuboNcomponentionicfluid.h:22-65)
// ---------------------------------------------------------------------------
template <class READLOG, class TFLOAT = double, class TFLOAT_READ = double>
class GreenKuboNComponentIonicFluid
    : public VectorOp<GreenKuboNComponentIonicFluid<READLOG, TFLOAT, TFLOAT_READ>, TFLOAT> {
    using VOp = VectorOp<GreenKuboNComponentIonicFluid<READLOG, TFLOAT, TFLOAT_READ>, TFLOAT>;
    using VOp::data_length;
    using VOp::vdata;

   public:
    GreenKuboNComponentIonicFluid(READLOG *traiettoria, std::string log, unsigned int skip,
                                  std::vector<std::string> headers, bool dump = false,
                                  unsigned int lunghezza_funzione_max = 0, unsigned int nthreads = 0,
                                  bool subtract_mean = false, unsigned int start_mean = 100, unsigned int n_seg = 1,
                                  bool do_bench = false, unsigned int n_seg_start = 10, unsigned int n_seg_stop = 100,
                                  bool mfma = false)
        : traiettoria(traiettoria), log(std::move(log)), skip(skip), scrivi_file(dump), lmax(lunghezza_funzione_max),
          subtract_mean(subtract_mean), start_mean(start_mean), n_seg(n_seg < 1 ? 1 : n_seg), mfma(mfma) {
        (void)nthreads;
        (void)do_bench;
        (void)n_seg_start;
        (void)n_seg_stop;
        for (const auto &h : headers) {  // :62-69
            auto res = traiettoria->get_index_of(h);
            if (res.second) idx_j.push_back(res.first);
            else throw std::runtime_error("Errore: header '" + h + "' not found!\n");
        }
        N_corr = (unsigned int)(idx_j.size() * idx_j.size());
        narr = 3 * N_corr + 2;
        unsigned int c = 0;  // :78-102
        std::stringstream d;
        for (unsigned int i = 0; i < idx_j.size(); i++)
            for (unsigned int j = 0; j < idx_j.size(); j++)
                d << "#" << ++c << ": c(" << i << ", " << j << ")\n#" << ++c << ": var[c(" << i << ", " << j << ")]\n";
        for (unsigned int i = 0; i < idx_j.size(); i++)
            for (unsigned int j = 0; j < idx_j.size(); j++) {
                d << "#" << ++c << ": L(" << i << ", " << j << ")\n#" << ++c << ": var[L(" << i << ", " << j << ")]\n";
                d << "#" << ++c << ": L*(" << i << ", " << j << ")\n#" << ++c << ": var[L*(" << i << ", " << j
                  << ")]\n";
            }
        d << "#" << ++c << ": k_gk\n#" << ++c << ": var[k_gk]\n";
        d << "#" << ++c << ": k_einst\n#" << ++c << ": var[k_einst]\n";
        c_descr = d.str();
    }
    unsigned int nExtraTimesteps(unsigned int n_b) {  // :117-119
        unsigned int v = traiettoria->n_timestep() / (n_b + 1) + 1;
        return (v < lmax || lmax == 0) ? v : lmax;
    }
    void reset(unsigned int n) {  // :122-131
        leff = (n < lmax || lmax == 0) ? n : lmax;
        data_length = leff * narr;
        ntimesteps = n;
        delete[] vdata;
        vdata = new TFLOAT[data_length];
        for (unsigned int i = 0; i < data_length; i++) vdata[i] = 0.0;
    }
    void calculate(unsigned int primo) {  // :156-380
        if (leff + primo + ntimesteps > traiettoria->n_timestep())
            throw std::runtime_error(
                "trajectory is too short for this kind of calculation. Select a different starting timestep or lower "
                "the size of the average or the lenght of the correlation function");
        // mfma: lag products as split-BF16 MFMA products (C(l) within ~1e-6 of
        // its scale); otherwise the FP64 form, the reference's values.
        anl_check((mfma ? anl_gk_calculate_mfma : anl_gk_calculate)(traiettoria->device(), idx_j.data(),
                                                                    (uint32_t)idx_j.size(), skip, lmax, ntimesteps,
                                                                    n_seg, subtract_mean, start_mean, primo, vdata));
        if (scrivi_file) {
            std::ofstream outfile(log + ".greekdump", std::ios::app);
            for (unsigned int it = 0; it < leff; it++) {
                for (unsigned int j = 0; j < narr; j++) outfile << vdata[it * narr + j] << " ";
                outfile << "\n";
            }
            outfile << "\n\n";
        }
    }
    GreenKuboNComponentIonicFluid &operator=(const GreenKuboNComponentIonicFluid &destra) {
        VOp::operator=(destra);
        return *this;
    }
    unsigned int get_narr() { return narr; }
    unsigned int get_indexOfKappa() { return 3 * N_corr; }
    std::string get_columns_description() { return c_descr; }
    std::vector<ssize_t> get_shape() { return {(ssize_t)leff, (ssize_t)narr}; }
    std::vector<ssize_t> get_stride() { return {(ssize_t)(narr * sizeof(TFLOAT)), (ssize_t)sizeof(TFLOAT)}; }

   private:
    READLOG *traiettoria;
    std::string log, c_descr;
    unsigned int skip, narr = 0, N_corr = 0;
    bool scrivi_file;
    unsigned int lmax, leff = 0, ntimesteps = 0;
    bool subtract_mean;
    unsigned int start_mean, n_seg;
    bool mfma;
    std::vector<uint32_t> idx_j;
};

// ---------------------------------------------------------------------------
// MediaVar / MediaVarCovar (calcoliblocchi.h:21-202)
// ---------------------------------------------------------------------------
// MediaVar: running mean and variance over blocks, element by element.  For
// block b (0-based) and every element k of the block's result x:
//     d = x[k] - m[k];   m[k] += d / (b + 1);   v[k] += (x[k] - m[k]) * d
// and after the last of n blocks v[k] /= (n - 1) n.  Per element these are
// the operations of the reference's whole-vector sequence
// (calcoliblocchi.h:21-65), in its order and precision, so the mean and the
// variance are bit-identical to it.
template <class T>
class MediaVar {
   public:
    MediaVar(T *mean, T *var) : mean_(mean), var_(var) {}
    void calcola_begin(unsigned int s, T *) {
        for (T *acc : {mean_, var_}) {
            acc->reset(s);
            acc->azzera();
        }
        seen_ = 0;
    }
    void calculate(T *block) {
        const unsigned int len = block->lunghezza();
        if (mean_->lunghezza() != len || var_->lunghezza() != len)
            throw std::runtime_error("Trying to operate on VectorOp of different sizes!");
        auto *x = block->access_vdata();
        auto *m = mean_->access_vdata();
        auto *v = var_->access_vdata();
        using F = std::remove_pointer_t<decltype(x)>;
        const F weight = (F)(seen_ + 1);
        for (unsigned int k = 0; k < len; k++) {
            const F d = x[k] - m[k];
            const F step = d / weight;
            m[k] += step;
            const F spread = x[k] - m[k];
            v[k] += spread * d;
        }
        seen_++;
    }
    void calcola_end(unsigned int n_b) {
        using F = std::remove_pointer_t<decltype(var_->access_vdata())>;
        *var_ /= (F)((n_b - 1) * n_b);
    }

   private:
    T *mean_, *var_;
    unsigned int seen_ = 0;
};

template <class T>
class MediaVarCovar {
   public:
    MediaVarCovar(unsigned int n_columns, std::vector<std::pair<unsigned int, unsigned int>> cvar)
        : n_columns(n_columns), cvar(std::move(cvar)) {
        for (auto &p : this->cvar)
            if (p.first >= n_columns || p.second >= n_columns)
                throw std::runtime_error("Errore: indice della covarianza fuori dal range!");
    }
    void calcola_begin(unsigned int, T *c) {
        size_block = c->lunghezza() / n_columns;
        iblock = 0;
        medie.assign(n_columns, std::vector<double>(size_block, 0.0));
        varianze.assign(n_columns, std::vector<double>(size_block, 0.0));
        covarianze.assign(cvar.size(), std::vector<double>(size_block, 0.0));
        dx.assign(n_columns, 0.0);
        dx2.assign(n_columns, 0.0);
    }
    void calculate(T *calcolo) {
        iblock++;
        for (unsigned int ib = 0; ib < size_block; ib++) {
            for (unsigned int ic = 0; ic < n_columns; ic++) {
                double x = calcolo->elemento(ib * n_columns + ic);
                dx[ic] = x - medie[ic][ib];
                medie[ic][ib] += dx[ic] / iblock;
                dx2[ic] = x - medie[ic][ib];
                varianze[ic][ib] += dx[ic] * dx2[ic];
            }
            for (unsigned int k = 0; k < cvar.size(); k++) covarianze[k][ib] += dx[cvar[k].first] * dx2[cvar[k].second];
        }
    }
    void calcola_end(unsigned int) {
        for (unsigned int ib = 0; ib < size_block; ib++) {
            for (unsigned int ic = 0; ic < n_columns; ic++) varianze[ic][ib] /= (iblock * (iblock - 1));
            for (unsigned int k = 0; k < cvar.size(); k++) covarianze[k][ib] /= (iblock * (iblock - 1));
        }
    }
    double *media(unsigned int c) { return medie.at(c).data(); }
    double *varianza(unsigned int c) { return varianze.at(c).data(); }
    double *covarianza(unsigned int c) { return covarianze.at(c).data(); }
    unsigned int size() { return size_block; }
    unsigned int n_column() { return (unsigned int)medie.size(); }
    unsigned int n_cvar() { return (unsigned int)covarianze.size(); }

   private:
    unsigned int n_columns, size_block = 0, iblock = 0;
    std::vector<std::pair<unsigned int, unsigned int>> cvar;
    std::vector<std::vector<double>> medie, varianze, covarianze;
    std::vector<double> dx, dx2;
};

// ---------------------------------------------------------------------------
// BlockAverageG (blockaverage.h:83-218), single process
// ---------------------------------------------------------------------------
template <class X, class = void>
struct counts_lines : std::false_type {};  // a log (n_timestep) rather than a trajectory
template <class X>
struct counts_lines<X, std::void_t<decltype(std::declval<X &>().n_timestep())>> : std::true_type {};
template <class TR>
inline unsigned int n_timesteps_of(TR *t) {
    if constexpr (counts_lines<TR>::value)
        return (unsigned int)t->n_timestep();
    else
        return (unsigned int)t->get_ntimesteps();
}

// BlockAverageG: the trajectory (or log) is cut into n_b consecutive blocks
// of equal length, chosen so that the last block still has the lags its
// calculation reads past its end; one calculation object is reset and run on
// each block in turn and an accumulator (MediaVar by default) folds the
// block results.  Over a windowed reader the window is sized once to a block
// plus its lags and moved to each block before its calculation; a reader
// that can read ahead starts on the following block's window meanwhile.
template <class TR, class T, typename... Args>
class BlockAverageG {
    // readers that can read the next block's window ahead
    template <class X, class = void>
    struct reads_ahead : std::false_type {};
    template <class X>
    struct reads_ahead<X, std::void_t<decltype(std::declval<X &>().set_next_window(size_t()))>> : std::true_type {};
    // readers that hold only a window of timesteps (blockaverage.h:32-51:
    // the LAMMPS binary; logs and in-memory arrays hold everything)
    template <class X, class = void>
    struct windowed : std::false_type {};
    template <class X>
    struct windowed<X, std::void_t<decltype(std::declval<X &>().set_access_at(size_t())),
                                   decltype(std::declval<X &>().set_data_access_block_size(size_t()))>>
        : std::true_type {};

    // Block length: the timesteps left after the lags of the whole run,
    // shared evenly (blockaverage.h:113-121; the reference aborts where this
    // throws).
    unsigned int block_length(T &probe) const {
        const long total = (long)n_timesteps_of(source_);
        const long extra = (long)probe.nExtraTimesteps(n_blocks_);
        const long len = (total - extra) / (long)n_blocks_;
        if (len <= 0) throw std::runtime_error("Cannot divide the trajectory in the requested blocks");
        return (unsigned int)len;
    }

   public:
    BlockAverageG(TR *t, const unsigned int &numero_blocchi) : source_(t), n_blocks_(numero_blocchi) {}

    // blockaverage.h:107-190 (one process: MPI's round-robin over blocks is
    // not part of this library)
    template <class Calcolo>
    void calcola_custom(Calcolo *acc, Args... arg) {
        work_ = std::make_unique<T>(source_, arg...);
        const unsigned int len = block_length(*work_);
        work_->reset(len);
        acc->calcola_begin(len, work_.get());
        if constexpr (windowed<TR>::value)
            source_->set_data_access_block_size((size_t)len + work_->nExtraTimesteps(n_blocks_));
        ready_ = true;
        for (unsigned int b = 0; b < n_blocks_; b++) {
            const size_t first = (size_t)b * len;
            work_->reset(len);
            if constexpr (windowed<TR>::value) source_->set_access_at(first);
            if constexpr (reads_ahead<TR>::value)
                if (b + 1 < n_blocks_) source_->set_next_window(first + len);
            work_->calculate(first);
            acc->calculate(work_.get());
        }
        acc->calcola_end(n_blocks_);
    }
    // blockaverage.h:194-203: mean and variance over the blocks
    void calculate(Args... arg) {
        mean_ = std::make_unique<T>(source_, arg...);
        var_ = std::make_unique<T>(source_, arg...);
        MediaVar<T> folded(mean_.get(), var_.get());
        calcola_custom(&folded, arg...);
    }
    T *media() { return checked(mean_.get()); }
    T *varianza() { return checked(var_.get()); }
    T *puntatoreCalcolo() { return checked(work_.get()); }

   private:
    T *checked(T *p) const {
        if (!ready_ || !p) throw std::runtime_error("block average not computed");
        return p;
    }
    TR *source_;
    unsigned int n_blocks_;
    std::unique_ptr<T> work_, mean_, var_;
    bool ready_ = false;
};

}  // namespace analisi
