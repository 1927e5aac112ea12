// ReadLog<TFLOAT>: the column-formatted text log reader of rikigigi/analisi
// v0.5.2 (readlog.{h,cpp}), with its "#traj:JZ N q1 ... qN" columns: a
// charge current computed from the per-type centre-of-mass velocities of a
// LAMMPS-binary Trajectory (readlog.cpp:207-293).  Same constructor, header
// and line rules, column indexing and error messages; device() mirrors the
// table on the GPU for GreenKuboNComponentIonicFluid<ReadLog<>>.
#pragma once

#include <cerrno>
#include <cstdlib>
#include <iterator>
#include <limits>

#include "lammps_traj.h"

namespace analisi {

template <class TFLOAT = double>
class ReadLog {
   public:
    // readlog.cpp:26-205.  The header is the line just before the first line
    // made only of number characters; every later such line whose fields all
    // parse is a row (every skip-th one is kept); a "Step" column becomes the
    // row's timestep and is not stored.
    ReadLog(const std::string &filename, Trajectory *t = nullptr, unsigned int skip = 1, unsigned int nthreads = 0,
            unsigned int nbatch = 200, std::vector<std::string> req_headers = std::vector<std::string>())
        : traiettoria(t), skip(skip == 0 ? 1 : skip), nthreads(nthreads), nbatch(nbatch) {
        std::ifstream log(filename);
        std::string tmp, header;
        step_index = std::numeric_limits<unsigned int>::max();
        while (log.good()) {
            header = tmp;
            std::getline(log, tmp);
            if (tmp.length() > 1 && tmp.back() == '\r') tmp.pop_back();
            if (if_only_numbers(tmp) && tmp.length() > 1) break;
        }
        if (header.size() == 0) throw std::runtime_error("Cannot find the column headers in file\"" + filename + "\"\n");
        // column names: the tokens between blanks and tabs (readlog.cpp:62-79)
        const std::string delim = " \t";
        size_t start = 0, end = header.find_first_of(delim);
        while (end != std::string::npos) {
            if (end != start && start < header.length()) headers.push_back(header.substr(start, end - start));
            start = end + 1;
            end = header.find_first_of(delim, start);
            if (!headers.empty() && headers.back() == "Step") step_index = (unsigned int)headers.size() - 1;
        }
        if (end != start && start < header.length()) headers.push_back(header.substr(start, end - start));
        if (headers.empty()) throw std::runtime_error("Cannot find the column headers in file\"" + filename + "\"\n");
        if (headers.back() == "Step") step_index = (unsigned int)headers.size() - 1;

        unsigned int n_columns_from_binary = 0;
        data_size_from_binary = 0;
        for (const auto &h : req_headers) {
            if (h == "Step") throw std::runtime_error("\"Step\" columns header name is reserved for step number");
            auto Qs = qs(h);
            if (Qs.first != "") {
                n_columns_from_binary++;
                q_current_type.push_back(Qs);
            }
        }
        data_size_from_binary = n_columns_from_binary * 3;
        const bool has_step = step_index != std::numeric_limits<unsigned int>::max();
        data_size = (has_step ? headers.size() - 1 : headers.size()) + data_size_from_binary;

        unsigned int cont = 0;
        std::vector<TFLOAT> row(data_size, 0.0);
        auto take_line = [&](const std::string &l) {
            if (!(if_only_numbers(l) && l.length() > 1)) return;
            std::stringstream sstr(l);
            unsigned int ts = 0;
            for (unsigned int i = 0; i < headers.size(); i++) {
                if (i < step_index) sstr >> row[i];
                else if (i > step_index) sstr >> row[i - 1];
                else sstr >> ts;
            }
            if (sstr.fail()) return;
            if (cont++ % this->skip == 0) {
                data.insert(data.end(), row.begin(), row.end());
                timesteps.push_back(ts > 0 ? ts : cont);
            }
        };
        take_line(tmp);
        while (log.good()) {
            std::getline(log, tmp);
            if (!log.good() && tmp.empty()) break;
            take_line(tmp);
        }
        if (data.empty())
            std::cerr << "Warning: I was not able to read anything (does the header has the same number of columns of "
                         "the data?)\n";
    }
    ~ReadLog() {
        if (dev) anl_series_destroy(dev);
    }
    ReadLog(const ReadLog &) = delete;
    ReadLog &operator=(const ReadLog &) = delete;

    TFLOAT *line(unsigned int index) { return &data[(size_t)index * data_size]; }
    unsigned int n_timestep() { return (unsigned int)(data.size() / data_size); }
    unsigned int n_data() { return (unsigned int)data_size; }
    unsigned int timestep(unsigned int index) { return timesteps.at(index); }

    // readlog.cpp:296-325: index of a column among the non-Step columns, or
    // of a "#traj:" current's first component
    std::pair<unsigned int, bool> get_index_of(const std::string &header) {
        if (header.size() == 0) throw std::runtime_error("Error: zero sized header");
        unsigned int idx = 0;
        if (header[0] != '#') {
            for (unsigned int i = 0; i < headers.size(); i++) {
                if (headers.at(i) != "Step") {
                    if (headers.at(i) == header) return {idx, true};
                    idx++;
                }
            }
            return {idx, false};
        }
        idx = get_calc_j_index(header);
        return {(unsigned int)(data_size - data_size_from_binary + idx * 3), true};
    }
    int need_binary(const std::vector<std::string> &hs) {
        int res = 0;
        for (const auto &h : hs)
            if (qs(h).first != "") res++;
        return res;
    }
    // readlog.cpp:253-293: every "#traj:" current, per row ts, is
    // sum over types of v_cm(ts, type) q_type, read from t in n_b windows
    void calc_currents(Trajectory *t, unsigned int n_b) {
        traiettoria = t;
        const unsigned int timesteps_tot = n_timestep();
        const unsigned int n_data_b = timesteps_tot / n_b;
        const size_t ntypes = t->get_ntypes();
        for (const auto &q : q_current_type)
            if (q.second.size() != ntypes)
                throw std::runtime_error(
                    "Cannot calculate current: number of coefficients must be equal to the number of atomic types in "
                    "the trajectory!\n");
        t->set_data_access_block_size(n_data_b);
        for (unsigned int ib = 0; ib < n_b; ib++) {
            unsigned int ultimo = (ib + 1) * n_data_b;
            if (ib == n_b - 1) {
                ultimo = timesteps_tot;
                t->set_data_access_block_size(ultimo - ib * n_data_b);
            }
            t->set_access_at(n_data_b * ib);
            for (unsigned int ts = ib * n_data_b; ts < ultimo; ts++)
                for (size_t i = 0; i < q_current_type.size(); i++) {
                    const double *v_cm = t->velocity_cm(ts, 0);
                    TFLOAT *dst = &data[(size_t)ts * data_size + (data_size - data_size_from_binary) + i * 3];
                    for (int c = 0; c < 3; c++) dst[c] = 0.0;
                    for (size_t icm = 0; icm < ntypes; icm++)
                        for (int c = 0; c < 3; c++) dst[c] += v_cm[icm * 3 + c] * q_current_type[i].second[icm];
                }
        }
        if (dev) {  // the table changed: the device copy is rebuilt on next use
            anl_series_destroy(dev);
            dev = nullptr;
        }
    }

    anl_series *device() {
        if (!dev) anl_check(anl_series_create(data.data(), n_timestep(), (uint32_t)data_size, &dev));
        return dev;
    }

   private:
    static bool if_only_numbers(const std::string &str) {
        return str.find_first_not_of("Ee0123456789 \t.-+\r\n") == std::string::npos;
    }
    // readlog.cpp:223-249: "#traj:JZ N q1 ... qN" -> (header, charges)
    static std::pair<std::string, std::vector<TFLOAT>> qs(const std::string &header) {
        if (header.size() == 0 || header[0] != '#') return {};
        std::stringstream iss(header);
        std::vector<std::string> tok{std::istream_iterator<std::string>{iss}, std::istream_iterator<std::string>{}};
        const std::string msg = "Error in header format '" + header + "'. It must be in the form '#traj:JZ N q1 ... qN'.\n";
        if (tok.size() < 3) throw std::runtime_error(msg);
        errno = 0;
        const long n = std::strtol(tok.at(1).c_str(), nullptr, 0);
        if (n <= 0 || errno == ERANGE || tok.size() != (size_t)n + 2) throw std::runtime_error(msg);
        std::vector<TFLOAT> c((size_t)n);
        for (long i = 2; i < n + 2; i++) {
            c[(size_t)(i - 2)] = (TFLOAT)std::strtold(tok.at((size_t)i).c_str(), nullptr);
            if (errno == ERANGE) throw std::runtime_error(msg);
        }
        return {header, c};
    }
    unsigned int get_calc_j_index(const std::string &header) {
        for (unsigned int i = 0; i < q_current_type.size(); i++)
            if (q_current_type[i].first == header) return i;
        std::cerr << "Error: cannot find the header \"" << header << "\" in the list '";
        for (const auto &q : q_current_type) std::cerr << " '" << q.first << "'";
        std::cerr << "\n";
        throw std::runtime_error("Wrong header name");
    }

    Trajectory *traiettoria;
    std::vector<std::string> headers;
    std::vector<TFLOAT> data;
    std::vector<unsigned int> timesteps;
    unsigned int skip, nthreads, nbatch, step_index;
    size_t data_size = 0, data_size_from_binary = 0;
    std::vector<std::pair<std::string, std::vector<TFLOAT>>> q_current_type;
    anl_series *dev = nullptr;
};

}  // namespace analisi
