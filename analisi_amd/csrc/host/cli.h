// The analisi command line's block-averaged text outputs (analisi/main.cpp),
// for the three paths of the pair-correlation family: Green-Kubo (-H), mean
// square displacement (-q, -Q, --mean-square-displacement-self) and g(r,t)
// (-g).  Each runs the same BlockAverageG loop as the command line over the
// device path and writes what the command line writes to stdout,
// with an ostream's default formatting (6 significant digits), so the text
// can be compared with the command line's own output files.
#pragma once

#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "analisi.h"

namespace analisi {
namespace cli {

template <class X, class = void>
struct has_pbc_wrap : std::false_type {};
template <class X>
struct has_pbc_wrap<X, std::void_t<decltype(std::declval<X &>().set_pbc_wrap(true))>> : std::true_type {};

// main.cpp:388-513 with the default factors (1, 1) or the two of -F: the
// "#Factors" line, the factor header, the columns' legend, then per lag the
// mean and variance of every column times its factor (factor_conv *
// factor_intToCorr for the n^2 correlations, factor_conv for the rest) and
// the covariance of every column pair of -z (cvar_list: pairs of column
// indices, main.cpp:266-274) times both factors.
template <class RL>
void green_kubo(std::ostream &o, RL &log, const std::vector<std::string> &headers, unsigned int blocknumber,
                unsigned int skip, unsigned int stop_acf, bool sub_mean, unsigned int sub_mean_start,
                unsigned int n_seg, double factor_conv = 1.0, double factor_intToCorr = 1.0, bool mfma = false,
                const std::vector<unsigned int> &cvar_list = {}) {
    using GK = GreenKuboNComponentIonicFluid<RL, double, double>;
    if (cvar_list.size() % 2 != 0)
        throw std::runtime_error("Error: covariance indexes list must contain an even number of elements\n");
    std::vector<std::pair<unsigned int, unsigned int>> cvar;
    for (size_t i = 0; i + 1 < cvar_list.size(); i += 2) cvar.emplace_back(cvar_list[i], cvar_list[i + 1]);
    o << "#Factors: " << factor_conv << ", " << factor_conv * factor_intToCorr << "\n";
    BlockAverageG<RL, GK, std::string, unsigned int, std::vector<std::string>, bool, unsigned int, unsigned int, bool,
                  unsigned int, unsigned int, bool, unsigned int, unsigned int, bool>
        ba(&log, blocknumber);
    const unsigned int nh = (unsigned int)headers.size();
    const unsigned int narr = nh * nh * 3 + 2;
    MediaVarCovar<GK> mvc(narr, cvar);
    ba.calcola_custom(&mvc, std::string(""), skip, headers, false, stop_acf, 1u, sub_mean, sub_mean_start, n_seg,
                      false, 10u, 100u, mfma);
    std::vector<double> f(narr, factor_conv);
    for (unsigned int j = 0; j < nh * nh; j++) f[j] = factor_conv * factor_intToCorr;
    o << "# factor integral, factor correlation\n#" << factor_conv << " " << factor_intToCorr << "\n";
    o << ba.puntatoreCalcolo()->get_columns_description();
    for (unsigned int i = 0; i < mvc.size(); i++) {
        for (unsigned int j = 0; j < narr; j++)
            o << mvc.media(j)[i] * f[j] << " " << mvc.varianza(j)[i] * f[j] * f[j] << " ";
        for (unsigned int j = 0; j < mvc.n_cvar(); j++)
            o << mvc.covarianza(j)[i] * f[cvar[j].first] * f[cvar[j].second] << " ";
        o << "\n";
    }
}

// main.cpp:520-551: one line per lag, the mean and variance of each type's
// MSD (and, with the centre-of-mass option, then each type's centre-of-mass
// MSD), every value followed by a space.
template <class TR>
void mean_square_displacement(std::ostream &o, TR &t, unsigned int blocknumber, unsigned int skip,
                              unsigned int stop_acf, bool msd_cm, bool msd_self) {
    BlockAverageG<TR, MSD<TR>, unsigned int, unsigned int, unsigned int, bool, bool, bool> ba(&t, blocknumber);
    ba.calculate(skip, stop_acf, 1u, msd_cm, msd_self, false);
    const unsigned int w = (unsigned int)t.get_ntypes() * (msd_cm ? 2u : 1u);
    MSD<TR> *m = ba.media(), *v = ba.varianza();
    for (unsigned int i = 0; i < m->lunghezza() / w; i++) {
        for (unsigned int j = 0; j < w; j++) o << m->elemento(i * w + j) << " " << v->elemento(i * w + j) << " ";
        o << "\n";
    }
}

// main.cpp:552-581: the columns' legend, then for every lag t (step every)
// and bin r the line "t r" followed by the mean and variance of each channel,
// and two empty lines after each lag.  The command line wraps positions into
// the cell first (set_pbc_wrap(true) on its LAMMPS trajectory); a trajectory
// type with that switch gets it set here.
template <class TR>
void gofrt(std::ostream &o, TR &tr, unsigned int blocknumber, double rmin, double rmax, unsigned int nbin,
           unsigned int stop_acf, unsigned int skip, unsigned int every) {
    if constexpr (has_pbc_wrap<TR>::value) tr.set_pbc_wrap(true);
    using G = Gofrt<double, TR>;
    BlockAverageG<TR, G, double, double, unsigned int, unsigned int, unsigned int, unsigned int, unsigned int, bool,
                  int>
        ba(&tr, blocknumber);
    ba.calculate(rmin, rmax, nbin, stop_acf, 1u, skip, every, false, (int)ANL_GOFRT_AUTO);
    const unsigned int nch = (unsigned int)(tr.get_ntypes() * (tr.get_ntypes() + 1));
    const unsigned int tmax = ba.media()->lunghezza() / nbin / nch;
    if (every == 0) every = 1;
    o << ba.puntatoreCalcolo()->get_columns_description();
    for (unsigned int t = 0; t < tmax; t += every) {
        for (unsigned int r = 0; r < nbin; r++) {
            o << t << " " << r;
            for (unsigned int c = 0; c < nch; c++) {
                const unsigned int k = t * nch * nbin + nbin * c + r;
                o << " " << ba.media()->elemento(k) << " " << ba.varianza()->elemento(k);
            }
            o << "\n";
        }
        o << "\n\n";
    }
}

}  // namespace cli
}  // namespace analisi
