// pybind11 module with the pyanalisi classes of the pair-correlation path
// (pyanalisi/src/pyanalisi.cpp:65-153, :241-279, :281-373, :377-450,
// :474-550 of rikigigi/analisi v0.5.2): Trajectory, Traj (LAMMPS binary),
// ReadLog, ReadLog_file (text), BoxFormat, Gofrt / Gofrt_lammps,
// MeanSquareDisplacement / MeanSquareDisplacement_lammps, GreenKubo /
// GreenKubo_file, info(), has_mmap().  Same
// constructor arguments, methods and buffer protocol (FP64, zero copy on
// vdata); the calculations run on the GPU through the C ABI.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "analisi.h"
#include "cli.h"
#include "lammps_traj.h"
#include "readlog.h"

namespace py = pybind11;
using namespace analisi;

namespace {

template <typename T>
bool has_no_stride(const py::buffer_info &buf) {  // buffer_utils.h:7-15
    ssize_t stride = sizeof(T);
    for (int i = (int)buf.ndim - 1; i >= 0; --i) {
        if (buf.strides[i] != stride) return false;
        stride *= buf.shape[i];
    }
    return true;
}

// trajectory_numpy.cpp:7-80 argument checks, same messages
Trajectory_numpy *make_trajectory(py::buffer bpos, py::buffer bvel, py::buffer btypes, py::buffer bbox,
                                  Trajectory_numpy::BoxFormat fmt, bool pbc_wrap, bool save_rot) {
    py::buffer_info ip = bpos.request(), iv = bvel.request(), it = btypes.request(), ib = bbox.request();
    if (ip.ndim != 3) throw std::runtime_error("Wrong number of dimension of position array (must be 3)");
    if (iv.ndim != 3) throw std::runtime_error("Wrong number of dimension of velocities array (must be 3)");
    if (ip.shape[2] != 3)
        throw std::runtime_error("Wrong number of cartesian components in the third dimension of positions array");
    if (iv.shape[2] != 3)
        throw std::runtime_error("Wrong number of cartesian components in the third dimension of velocities array");
    for (size_t i = 0; i < ip.shape.size(); ++i)
        if (ip.shape[i] != iv.shape[i]) throw std::runtime_error("Shape of positions and velocities array is different");
    if (it.ndim != 1) throw std::runtime_error("Wrong number of dimension of types array (must be 1)");
    if (it.shape[0] != ip.shape[1]) throw std::runtime_error("Wrong size of the type array");
    if (fmt == Trajectory_numpy::BoxFormat::Cell_vectors) {
        if (ib.ndim != 3)
            throw std::runtime_error("Wrong number of dimensions of box array (must be 3) for cell matrix format");
        if (ib.shape[0] != ip.shape[0] || ib.shape[1] != 3 || ib.shape[2] != 3)
            throw std::runtime_error("Wrong shape of box array: must be (nsteps, 3, 3) for matrix format");
    } else {
        if (ib.ndim != 2)
            throw std::runtime_error(
                "Wrong number of dimensions of box array (must be 2 for lammps ortho/triclinic cell format)");
        if (ib.shape[0] != ip.shape[0])
            throw std::runtime_error("Wrong shape of box array: first dimension must be nsteps");
        if (fmt == Trajectory_numpy::BoxFormat::Lammps_ortho && ib.shape[1] != 6)
            throw std::runtime_error("Wrong shape of box array: must be (:, 6) for orthogonal lammps cell format");
        if (fmt == Trajectory_numpy::BoxFormat::Lammps_triclinic && ib.shape[1] != 9)
            throw std::runtime_error("Wrong shape of box array: must be (:, 9) for triclinic lammps cell format");
    }
    if (ib.format != py::format_descriptor<double>::format())
        throw std::runtime_error("Format of box array should be double");
    if (it.format != py::format_descriptor<int>::format())
        throw std::runtime_error("Format of types array should be int (" + py::format_descriptor<int>::format() +
                                 ") while it was " + it.format);
    if (iv.format != py::format_descriptor<double>::format())
        throw std::runtime_error("Format of velocities array should be double");
    if (ip.format != py::format_descriptor<double>::format())
        throw std::runtime_error("Format of positions array should be double");
    if (!has_no_stride<double>(ib)) throw std::runtime_error("Unsupported stride in box array");
    if (!has_no_stride<int>(it)) throw std::runtime_error("Unsupported stride in types array");
    if (!has_no_stride<double>(iv)) throw std::runtime_error("Unsupported stride in vel array");
    if (!has_no_stride<double>(ip)) throw std::runtime_error("Unsupported stride in pos array");
    return new Trajectory_numpy((const double *)ip.ptr, (const double *)iv.ptr, (const int *)it.ptr,
                                (const double *)ib.ptr, (size_t)ip.shape[0], (size_t)ip.shape[1], fmt, pbc_wrap,
                                save_rot);
}

template <class C>
py::buffer_info vbuffer(C &c) {
    return py::buffer_info(c.access_vdata(), sizeof(double), py::format_descriptor<double>::format(),
                           (ssize_t)c.get_shape().size(), c.get_shape(), c.get_stride());
}

py::array_t<double> copy_array(const double *p, std::vector<ssize_t> shape) {
    py::array_t<double> a(shape);
    std::memcpy(a.mutable_data(), p, a.size() * sizeof(double));
    return a;
}

// trajectory_common_interfaces (pyanalisi.cpp:281-373): the methods both
// trajectory classes share; the copies cover the loaded timesteps
template <class Tk, class C>
C &common_interfaces(C &c) {
    c.def("write_lammps_binary", &Tk::dump_lammps_bin_traj, py::arg("filename"), py::arg("start"),
          py::arg("stop"))
        .def("get_positions_copy",
             [](Tk &t) {
                 if (t.get_nloaded_timesteps() == 0) return py::array_t<double>();
                 return copy_array(t.positions(t.get_current_timestep(), 0),
                                   {(ssize_t)t.get_nloaded_timesteps(), (ssize_t)t.get_natoms(), 3});
             })
        .def("get_velocities_copy",
             [](Tk &t) {
                 if (t.get_nloaded_timesteps() == 0) return py::array_t<double>();
                 return copy_array(t.velocity(t.get_current_timestep(), 0),
                                   {(ssize_t)t.get_nloaded_timesteps(), (ssize_t)t.get_natoms(), 3});
             })
        .def("get_box_copy",
             [](Tk &t) {
                 if (t.get_nloaded_timesteps() == 0) return py::array_t<double>();
                 return copy_array(t.box(t.get_current_timestep()),
                                   {(ssize_t)t.get_nloaded_timesteps(), (ssize_t)t.get_box_stride()});
             })
        .def("get_type", &Tk::get_type, "get the type used for the internal representation")
        .def("get_nloaded_timesteps", &Tk::get_nloaded_timesteps)
        .def("getNtimesteps", &Tk::get_ntimesteps)
        .def("get_current_timestep", &Tk::get_current_timestep)
        .def("getWrapPbc", &Tk::get_pbc_wrap)
        .def("minImage", [](Tk &t, size_t i, size_t j, size_t it, size_t ij) {
            py::array_t<double> a(4);
            double *x = a.mutable_data();
            x[3] = t.d2_minImage(i, j, it, ij, x);
            return a;
        });
    return c;
}

// Steinhardt with its lmax fixed, one Python class per lmax as in the reference
template <class T, int L>
struct SteinhardtL : Steinhardt<T> {
    SteinhardtL(T *t, typename Steinhardt<T>::Rminmax_t rminmax, unsigned int nbin, unsigned int nbin_st,
                std::vector<unsigned int> l_hist, unsigned int nthreads, unsigned int skip, bool do_histogram,
                typename Steinhardt<T>::NeighListSpec nls, bool averaged)
        : Steinhardt<T>(t, rminmax, nbin, nbin_st, std::move(l_hist), nthreads, skip, do_histogram, nls, averaged,
                        (unsigned int)L) {}
};

template <class T, int L>
void define_steinhardt(py::module &m, const std::string &suffix) {
    using SL = SteinhardtL<T, L>;
    py::class_<SL>(m, ("SteinhardtOrderParameterHistogram_" + std::to_string(L) + suffix).c_str(),
                   py::buffer_protocol())
        .def(py::init<T *, typename SL::Rminmax_t, unsigned int, unsigned int, std::vector<unsigned int>, unsigned int,
                      unsigned int, bool, typename SL::NeighListSpec, bool>(),
             py::arg("trajectory"), py::arg("rminmax"), py::arg("nbin"), py::arg("histogram_size"),
             py::arg("l_list"), py::arg("nthreads") = 2, py::arg("skip") = 1, py::arg("do_histogram") = true,
             py::arg("neighList") = typename SL::NeighListSpec{}, py::arg("averaged_order") = false)
        .def("reset", &SL::reset)
        .def("calculate", &SL::calculate)
        .def_buffer([](SL &s) { return vbuffer(s); });
}

// define_atomic_traj (pyanalisi.cpp:65-153): the pair-correlation operators
// over one trajectory class, named with its suffix
template <class T>
void define_atomic_traj(py::module &m, const std::string &suffix) {
    using G = Gofrt<double, T>;
    using M = MSD<T>;
    py::class_<G>(m, ("Gofrt" + suffix).c_str(), py::buffer_protocol())
        .def(py::init<T *, double, double, unsigned int, unsigned int, unsigned int, unsigned int, unsigned int, bool,
                      int>(),
             py::arg("trajectory"), py::arg("rmin"), py::arg("rmax"), py::arg("nbin"), py::arg("tmax") = 0,
             py::arg("nthreads") = 0, py::arg("skip") = 1, py::arg("every") = 1, py::arg("debug") = false,
             py::arg("form") = ANL_GOFRT_AUTO)
        .def("reset", &G::reset)
        .def("getNumberOfExtraTimestepsNeeded", &G::nExtraTimesteps)
        .def("calculate", &G::calculate)
        .def("calc_single_th", &G::calc_single_th)
        .def("last_stats",
             [](G &g) {
                 int f = -1;
                 uint64_t c = 0, r = 0;
                 anl_check(anl_gofrt_last_stats(g.handle(), &f, &c, &r));
                 return py::make_tuple(f, c, r);
             })
        .def_buffer([](G &g) { return vbuffer(g); });

    py::class_<M>(m, ("MeanSquareDisplacement" + suffix).c_str(), py::buffer_protocol())
        .def(py::init<T *, unsigned int, unsigned int, unsigned int, bool, bool, bool>(), py::arg("trajectory"),
             py::arg("skip") = 1, py::arg("tmax") = 0, py::arg("nthreads") = 0, py::arg("cm") = false,
             py::arg("cm_self") = false, py::arg("debug") = false)
        .def("reset", &M::reset)
        .def("getNumberOfExtraTimestepsNeeded", &M::nExtraTimesteps)
        .def("calculate", &M::calculate)
        .def_buffer([](M &x) { return vbuffer(x); })
        .def("__enter__", [](M &x) -> M & { return x; })
        .def("__exit__", [](M &, py::object, py::object, py::object) {});

    // SphericalCorrelations (pyanalisi.cpp:85-111), lmax 10 by default
    using S = SphericalCorrelations<T>;
    py::class_<S>(m, ("SphericalCorrelations" + suffix).c_str(), py::buffer_protocol())
        .def(py::init<T *, typename S::rminmax_t, size_t, size_t, size_t, size_t, size_t, bool,
                      typename S::NeighListSpec, unsigned int>(),
             py::arg("trajectory"), py::arg("rminmax"), py::arg("nbin"), py::arg("tmax") = 0, py::arg("nthreads") = 0,
             py::arg("skip") = 1, py::arg("buffer_size") = 10, py::arg("debug") = false,
             py::arg("neighList") = typename S::NeighListSpec{}, py::arg("lmax") = (unsigned int)ANL_SPH_LMAX_MAX)
        .def("reset", &S::reset)
        .def("calculate", &S::calculate)
        .def("getNumberOfExtraTimestepsNeeded", &S::nExtraTimesteps)
        .def("get_columns_description", &S::get_columns_description)
        .def("calc",
             [](S &s, size_t timestep) {
                 std::vector<double> v = s.calc(timestep);
                 return copy_array(v.data(), {(ssize_t)v.size()});
             })
        .def_buffer([](S &s) { return vbuffer(s); });

    // SteinhardtOrderParameterHistogram_<l> (pyanalisi.cpp:22-60, 114-122), l = 2..10
    define_steinhardt<T, 2>(m, suffix);
    define_steinhardt<T, 3>(m, suffix);
    define_steinhardt<T, 4>(m, suffix);
    define_steinhardt<T, 5>(m, suffix);
    define_steinhardt<T, 6>(m, suffix);
    define_steinhardt<T, 7>(m, suffix);
    define_steinhardt<T, 8>(m, suffix);
    define_steinhardt<T, 9>(m, suffix);
    define_steinhardt<T, 10>(m, suffix);
}

// gk (pyanalisi.cpp:241-279): GreenKubo over one log class
template <class RLT>
void define_gk(py::module &m, const std::string &suffix) {
    using G = GreenKuboNComponentIonicFluid<RLT, double, double>;
    py::class_<G>(m, ("GreenKubo" + suffix).c_str(), py::buffer_protocol())
        .def(py::init<RLT *, std::string, unsigned int, std::vector<std::string>, bool, unsigned int, unsigned int,
                      bool, unsigned int, unsigned int, bool, unsigned int, unsigned int, bool>(),
             py::arg("log"), py::arg("dump"), py::arg("skip"), py::arg("headers"), py::arg("dump_files") = false,
             py::arg("lmax") = 0, py::arg("nthreads") = 0, py::arg("subtract_mean") = false,
             py::arg("start_mean") = 100, py::arg("n_seg") = 1, py::arg("do_bench") = false,
             py::arg("n_seg_start") = 10, py::arg("n_seg_stop") = 100, py::arg("mfma") = false)
        .def("getNumberOfExtraTimestepsNeeded", &G::nExtraTimesteps)
        .def("reset", &G::reset)
        .def("calculate", &G::calculate)
        .def("get_columns_description", &G::get_columns_description)
        .def_buffer([](G &g) { return vbuffer(g); });
}

// BlockAverageG + MediaVar of Gofrt / MSD (blockaverage.h:194-203), the
// CLI's -g and -Q paths: (mean, variance) in the calculation's layout.  Over
// a Traj, each block's calculation moves the window to the block (the window
// must hold a block and its lags).
template <class T>
void define_block_averages(py::module &m) {
    using G = Gofrt<double, T>;
    using M = MSD<T>;
    m.def("gofrt_block_average",
          [](T &t, unsigned int nblocks, double rmin, double rmax, unsigned int nbin, unsigned int tmax,
             unsigned int skip, unsigned int every) {
              BlockAverageG<T, G, double, double, unsigned int, unsigned int, unsigned int, unsigned int, unsigned int,
                            bool, int>
                  ba(&t, nblocks);
              ba.calculate(rmin, rmax, nbin, tmax, 1u, skip, every, false, (int)ANL_GOFRT_AUTO);
              auto shape = ba.media()->get_shape();
              return py::make_tuple(copy_array(ba.media()->access_vdata(), shape),
                                    copy_array(ba.varianza()->access_vdata(), shape));
          });
    m.def("msd_block_average", [](T &t, unsigned int nblocks, unsigned int skip, unsigned int tmax, bool cm,
                                  bool cm_self) {
        BlockAverageG<T, M, unsigned int, unsigned int, unsigned int, bool, bool, bool> ba(&t, nblocks);
        ba.calculate(skip, tmax, 1u, cm, cm_self, false);
        auto shape = ba.media()->get_shape();
        std::vector<ssize_t> sh(shape.begin(), shape.end());
        return py::make_tuple(copy_array(ba.media()->access_vdata(), sh), copy_array(ba.varianza()->access_vdata(), sh));
    });
    // the command line's stdout text of -q / -Q / --mean-square-displacement-self
    // (main.cpp:520-551) and of -g (main.cpp:552-581)
    m.def("cli_msd_text",
          [](T &t, unsigned int blocknumber, unsigned int skip, unsigned int stop_acf, bool msd_cm, bool msd_self) {
              std::ostringstream o;
              cli::mean_square_displacement(o, t, blocknumber, skip, stop_acf, msd_cm, msd_self);
              return o.str();
          },
          py::arg("traj"), py::arg("blocknumber") = 20, py::arg("skip") = 1, py::arg("stop_acf") = 0,
          py::arg("msd_cm") = false, py::arg("msd_self") = false);
    m.def("cli_gofrt_text",
          [](T &t, unsigned int blocknumber, double rmin, double rmax, unsigned int nbin, unsigned int stop_acf,
             unsigned int skip, unsigned int every) {
              std::ostringstream o;
              cli::gofrt(o, t, blocknumber, rmin, rmax, nbin, stop_acf, skip, every);
              return o.str();
          },
          py::arg("traj"), py::arg("blocknumber"), py::arg("rmin"), py::arg("rmax"), py::arg("nbin"),
          py::arg("stop_acf") = 0, py::arg("skip") = 1, py::arg("every") = 1);
}

// the command line's stdout text of -H (main.cpp:388-513)
template <class L>
void define_cli_gk(py::module &m) {
    m.def("cli_gk_text",
          [](L &log, std::vector<std::string> headers, unsigned int blocknumber, unsigned int skip,
             unsigned int stop_acf, bool sub_mean, unsigned int sub_mean_start, unsigned int n_seg, double factor_conv,
             double factor_intToCorr, bool mfma, std::vector<unsigned int> covariance) {
              std::ostringstream o;
              cli::green_kubo(o, log, headers, blocknumber, skip, stop_acf, sub_mean, sub_mean_start, n_seg,
                              factor_conv, factor_intToCorr, mfma, covariance);
              return o.str();
          },
          py::arg("log"), py::arg("headers"), py::arg("blocknumber") = 20, py::arg("skip") = 1,
          py::arg("stop_acf") = 0, py::arg("subtract_mean") = false, py::arg("start_mean") = 0, py::arg("n_seg") = 1,
          py::arg("factor_conv") = 1.0, py::arg("factor_intToCorr") = 1.0, py::arg("mfma") = false,
          py::arg("covariance") = std::vector<unsigned int>());
}

}  // namespace

using GofrtT = Gofrt<double, Trajectory_numpy>;
using MSDT = MSD<Trajectory_numpy>;
using RL = ReadLog_numpy<double>;
using GK = GreenKuboNComponentIonicFluid<RL, double, double>;

PYBIND11_MODULE(_pyanalisi, m) {
    m.doc() = "pyanalisi classes of the time-lagged pair-correlation path, computed on an MI355X";
    py::enum_<Trajectory_numpy::BoxFormat>(m, "BoxFormat", py::arithmetic())
        .value("Invalid", Trajectory_numpy::BoxFormat::Invalid)
        .value("CellVectors", Trajectory_numpy::BoxFormat::Cell_vectors)
        .value("LammpsOrtho", Trajectory_numpy::BoxFormat::Lammps_ortho)
        .value("LammpsTriclinic", Trajectory_numpy::BoxFormat::Lammps_triclinic);

    // Traj: the memory-mapped LAMMPS-binary trajectory (pyanalisi.cpp:377-450)
    py::class_<Trajectory> traj(m, "Traj", py::buffer_protocol());
    traj.def(py::init<std::string>(), py::arg("filename"))
        .def("setWrapPbc", &Trajectory::set_pbc_wrap)
        .def("setAccessWindowSize", [](Trajectory &t, int ts) { return (int)t.set_data_access_block_size((size_t)ts); })
        .def("setAccessStart", [](Trajectory &t, int ts) { return (int)t.set_access_at((size_t)ts); })
        .def("setPrefetch", &Trajectory::set_prefetch, py::arg("on"))
        .def("prefetch", [](Trajectory &t, size_t ts) { t.prefetch(ts); }, py::arg("timestep"))
        .def("serving_pos", &Trajectory::serving_pos)
        .def("toggle_pos_vel", &Trajectory::toggle_pos_vel)
        .def_buffer([](Trajectory &t) {
            return py::buffer_info(t.serving_pos() ? t.positions_data() : t.velocity_data(), sizeof(double),
                                   py::format_descriptor<double>::format(), 3, t.get_shape(), t.get_stride());
        })
        .def("get_lammps_id", [](Trajectory &t) { return py::array_t<int>(py::cast(t.get_lammps_id())); })
        .def("get_lammps_type", [](Trajectory &t) { return py::array_t<int>(py::cast(t.get_lammps_type())); })
        .def("get_timestep_lammps", &Trajectory::get_timestep_lammps);
    common_interfaces<Trajectory>(traj);

    py::class_<Trajectory_numpy> tnp(m, "Trajectory");
    tnp.def(py::init(&make_trajectory), py::keep_alive<1, 2>(), py::keep_alive<1, 3>(), py::keep_alive<1, 4>(),
            py::keep_alive<1, 5>(), py::arg("positions"), py::arg("velocities"), py::arg("types"),
            py::arg("lattice"), py::arg("format"), py::arg("pbc_wrap") = false,
            py::arg("save_rotation_matrix") = false)
        .def("get_rotation_matrix", [](Trajectory_numpy &t) {
            double *r = t.get_rotation_matrix(0);
            if (!r) return py::array_t<double>();
            return copy_array(r, {(ssize_t)t.get_ntimesteps(), 3, 3});
        });
    common_interfaces<Trajectory_numpy>(tnp);

    py::class_<RL>(m, "ReadLog")
        .def(py::init([](py::buffer data, std::vector<std::string> headers) {
                 py::buffer_info info = data.request();
                 if (info.ndim != 2) throw std::runtime_error("Wrong number of dimensions of the log array (must be 2)");
                 if (info.format != py::format_descriptor<double>::format())
                     throw std::runtime_error(
                         "Format of log array is wrong (should be some kind of float, try to change it if it does not "
                         "work)");
                 if (!has_no_stride<double>(info)) throw std::runtime_error("Unsupported stride in log array");
                 return new RL((const double *)info.ptr, (size_t)info.shape[0], (size_t)info.shape[1], headers);
             }),
             py::keep_alive<1, 2>())
        .def("getNtimesteps", &RL::n_timestep);

    define_atomic_traj<Trajectory_numpy>(m, "");
    define_atomic_traj<Trajectory>(m, "_lammps");

    // ReadLog_file: the text log reader (pyanalisi.cpp:455-471); calc_currents
    // fills its "#traj:" columns from a Traj, as the CLI does (main.cpp)
    using RLF = ReadLog<double>;
    py::class_<RLF>(m, "ReadLog_file")
        .def(py::init<std::string, Trajectory *, unsigned int, unsigned int, unsigned int, std::vector<std::string>>(),
             py::arg("filename"), py::arg("traj") = nullptr, py::arg("skip") = 1, py::arg("nthreads") = 0,
             py::arg("nbatch") = 200, py::arg("headers") = std::vector<std::string>())
        .def("getNtimesteps", &RLF::n_timestep)
        .def("calc_currents", &RLF::calc_currents, py::arg("traj"), py::arg("blocks"))
        .def("get_data_copy", [](RLF &l) {
            return copy_array(l.line(0), {(ssize_t)l.n_timestep(), (ssize_t)l.n_data()});
        })
        .def("timestep", &RLF::timestep);

    define_gk<RL>(m, "");
    define_gk<RLF>(m, "_file");

    // BlockAverageG + MediaVarCovar of GreenKubo, as the CLI's -H path runs it
    // (main.cpp:456-488): returns (mean, variance) as [narr][block size].
    m.def("gk_block_average",
          [](RL &log, std::vector<std::string> headers, unsigned int nblocks, unsigned int skip, unsigned int lmax,
             unsigned int n_seg, bool sub_mean, unsigned int start_mean, bool mfma) {
              BlockAverageG<RL, GK, std::string, unsigned int, std::vector<std::string>, bool, unsigned int,
                            unsigned int, bool, unsigned int, unsigned int, bool, unsigned int, unsigned int, bool>
                  ba(&log, nblocks);
              unsigned int narr = (unsigned int)(headers.size() * headers.size() * 3 + 2);
              MediaVarCovar<GK> mvc(narr, {});
              ba.calcola_custom(&mvc, std::string(""), skip, headers, false, lmax, 1u, sub_mean, start_mean, n_seg,
                                false, 10u, 100u, mfma);
              py::array_t<double> mean({(ssize_t)narr, (ssize_t)mvc.size()}), var({(ssize_t)narr, (ssize_t)mvc.size()});
              for (unsigned int c = 0; c < narr; c++) {
                  std::memcpy(mean.mutable_data() + (size_t)c * mvc.size(), mvc.media(c), mvc.size() * sizeof(double));
                  std::memcpy(var.mutable_data() + (size_t)c * mvc.size(), mvc.varianza(c), mvc.size() * sizeof(double));
              }
              return py::make_tuple(mean, var);
          },
          py::arg("log"), py::arg("headers"), py::arg("nblocks"), py::arg("skip"), py::arg("lmax"), py::arg("n_seg"),
          py::arg("subtract_mean"), py::arg("start_mean"), py::arg("mfma") = false);
    // BlockAverageG + MediaVar of Gofrt / MSD (blockaverage.h:194-203), the
    // CLI's -g and -Q paths: returns (mean, variance) in the calculation's layout.
    define_cli_gk<RL>(m);
    define_cli_gk<ReadLog<double>>(m);
    define_block_averages<Trajectory_numpy>(m);
    define_block_averages<Trajectory>(m);

    m.def("info", []() { return std::string(anl_info()); });
    m.def("has_mmap", []() { return true; });
}
