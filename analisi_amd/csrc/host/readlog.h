// ReadLog<TFLOAT>: the column-formatted text log of rikigigi/analisi v0.5.2
// (readlog.{h,cpp}) as a table of doubles for GreenKuboNComponentIonicFluid,
// plus "#traj:JZ N q1 ... qN" columns, charge currents built from the
// per-type centre-of-mass velocities of a LAMMPS-binary Trajectory
// (readlog.cpp:207-293).
//
// The file is read in three steps of this implementation's own:
//   1. scan_to_first_row() finds the first data row; the line before it is
//      the header (readlog.cpp:39-56);
//   2. a ColumnMap built from the header tells where each field of a row is
//      stored and which field is the timestep ("Step") (readlog.cpp:60-87);
//   3. rows are parsed in blocks of lines, each block split over threads;
//      the accepted rows are then appended in file order, every skip-th one
//      kept (readlog.cpp:114-185).
// A row is a line made only of number characters and blanks, longer than
// one character, whose fields all parse (std::istream extraction, field by
// field in file order).  The error messages are the reference's.
#pragma once

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <iterator>
#include <limits>
#include <thread>

#include "lammps_traj.h"

namespace analisi {

namespace logtext {

// readlog.cpp:327-329: the characters a data row may contain
inline bool number_chars_only(const std::string &s) {
    return s.find_first_not_of("Ee0123456789 \t.-+\r\n") == std::string::npos;
}
inline bool is_row_candidate(const std::string &s) { return s.length() > 1 && number_chars_only(s); }

// Non-empty runs of characters other than blank and tab.
inline std::vector<std::string> blank_tab_fields(const std::string &s) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (ch == ' ' || ch == '\t') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// Where the fields of a row go: the header's names in order; the last one
// called "Step" (if any) is the row's timestep and is not stored, the others
// are stored in header order.
struct ColumnMap {
    std::vector<std::string> names;
    long step = -1;
    size_t stored() const { return names.size() - (step >= 0 ? 1 : 0); }
    // index of a header name counted over the names other than "Step"
    // (readlog.cpp:306-316: every column called Step is passed over, which
    // is the stored index whenever Step appears at most once), or that count
    // if the name is absent
    size_t find(const std::string &name) const {
        size_t k = 0;
        for (const auto &n : names) {
            if (n == "Step") continue;
            if (n == name) return k;
            k++;
        }
        return k;
    }
};

inline ColumnMap map_columns(const std::string &header_line, const std::string &filename) {
    ColumnMap m;
    m.names = blank_tab_fields(header_line);
    if (m.names.empty()) throw std::runtime_error("Cannot find the column headers in file\"" + filename + "\"\n");
    for (size_t i = 0; i < m.names.size(); i++)
        if (m.names[i] == "Step") m.step = (long)i;
    return m;
}

// One "#traj:JZ N q1 ... qN" request: its text and the N per-type charges.
template <class TF>
struct ChargeCurrent {
    std::string spec;
    std::vector<TF> charge;
};

// readlog.cpp:223-249.  Returns false for a plain column name (no leading
// '#'); throws the reference's message for a malformed request.
template <class TF>
bool parse_charge_current(const std::string &spec, ChargeCurrent<TF> *out) {
    if (spec.empty() || spec[0] != '#') return false;
    const std::string msg = "Error in header format '" + spec + "'. It must be in the form '#traj:JZ N q1 ... qN'.\n";
    std::istringstream in(spec);
    const std::vector<std::string> words{std::istream_iterator<std::string>{in}, std::istream_iterator<std::string>{}};
    if (words.size() < 3) throw std::runtime_error(msg);
    errno = 0;
    const long count = std::strtol(words[1].c_str(), nullptr, 0);
    if (count <= 0 || errno == ERANGE || words.size() != (size_t)count + 2) throw std::runtime_error(msg);
    out->spec = spec;
    out->charge.clear();
    for (size_t w = 2; w < words.size(); w++) {
        out->charge.push_back((TF)std::strtold(words[w].c_str(), nullptr));
        if (errno == ERANGE) throw std::runtime_error(msg);
    }
    return true;
}

// Parses one candidate line: every header field in order, the Step field as
// an unsigned timestep, the others into dst[stored index].  False when the
// line is not a row (characters) or a field does not parse.
template <class TF>
bool parse_row(const std::string &line, const ColumnMap &cols, TF *dst, unsigned int *step_value) {
    if (!is_row_candidate(line)) return false;
    std::istringstream in(line);
    unsigned int ts = 0;
    size_t k = 0;
    for (size_t i = 0; i < cols.names.size(); i++) {
        if ((long)i == cols.step) in >> ts;
        else in >> dst[k++];
    }
    if (in.fail()) return false;
    *step_value = ts;
    return true;
}

}  // namespace logtext

template <class TFLOAT = double>
class ReadLog {
   public:
    // readlog.cpp:26-205.  nthreads threads parse blocks of nthreads*nbatch
    // lines; skip keeps every skip-th accepted row (0 is taken as 1).
    ReadLog(const std::string &filename, Trajectory *t = nullptr, unsigned int skip = 1, unsigned int nthreads = 0,
            unsigned int nbatch = 200, std::vector<std::string> req_headers = std::vector<std::string>())
        : traj_(t), keep_every_(skip == 0 ? 1 : skip) {
        std::ifstream in(filename);
        std::string first_row, header;
        scan_to_first_row(in, &header, &first_row);
        if (header.empty()) throw std::runtime_error("Cannot find the column headers in file\"" + filename + "\"\n");
        cols_ = logtext::map_columns(header, filename);
        for (const auto &h : req_headers) {
            if (h == "Step") throw std::runtime_error("\"Step\" columns header name is reserved for step number");
            logtext::ChargeCurrent<TFLOAT> c;
            if (logtext::parse_charge_current(h, &c)) currents_.push_back(c);
        }
        width_ = cols_.stored() + 3 * currents_.size();
        read_rows(in, first_row, std::max(1u, nthreads), std::max(1u, nbatch));
        if (table_.empty())
            std::cerr << "Warning: I was not able to read anything (does the header has the same number of columns of "
                         "the data?)\n";
    }
    ~ReadLog() {
        if (dev_) anl_series_destroy(dev_);
    }
    ReadLog(const ReadLog &) = delete;
    ReadLog &operator=(const ReadLog &) = delete;

    TFLOAT *line(unsigned int index) { return &table_[(size_t)index * width_]; }
    unsigned int n_timestep() { return (unsigned int)(table_.size() / width_); }
    unsigned int n_data() { return (unsigned int)width_; }
    unsigned int timestep(unsigned int index) { return step_of_row_.at(index); }

    // readlog.cpp:296-325: (column index, found) of a header name among the
    // stored columns, or of the first component of a "#traj:" current
    std::pair<unsigned int, bool> get_index_of(const std::string &header) {
        if (header.empty()) throw std::runtime_error("Error: zero sized header");
        if (header[0] == '#') return {(unsigned int)(cols_.stored() + 3 * current_index(header)), true};
        const size_t k = cols_.find(header);
        const bool found = header != "Step" && std::find(cols_.names.begin(), cols_.names.end(), header) !=
                                                   cols_.names.end();
        return {(unsigned int)k, found};
    }
    // number of "#traj:" requests among hs (readlog.cpp:251-262)
    int need_binary(const std::vector<std::string> &hs) {
        int n = 0;
        logtext::ChargeCurrent<TFLOAT> c;
        for (const auto &h : hs) n += logtext::parse_charge_current(h, &c) ? 1 : 0;
        return n;
    }
    // readlog.cpp:264-293: fills every "#traj:" current of every row ts with
    // sum over types of q_type v_cm(ts, type), the trajectory read in n_b
    // consecutive windows
    void calc_currents(Trajectory *t, unsigned int n_b) {
        traj_ = t;
        const size_t ntypes = t->get_ntypes();
        for (const auto &c : currents_)
            if (c.charge.size() != ntypes)
                throw std::runtime_error(
                    "Cannot calculate current: number of coefficients must be equal to the number of atomic types in "
                    "the trajectory!\n");
        if (n_b == 0) throw std::runtime_error("Cannot calculate current: zero blocks");
        const unsigned int rows = n_timestep();
        const unsigned int per_window = rows / n_b;
        t->set_data_access_block_size(per_window);
        for (unsigned int w = 0; w < n_b; w++) {
            const unsigned int first = w * per_window;
            const unsigned int end = (w + 1 == n_b) ? rows : first + per_window;
            if (w + 1 == n_b) t->set_data_access_block_size(end - first);
            t->set_access_at(first);
            for (unsigned int ts = first; ts < end; ts++) fill_currents_of_row(ts, t->velocity_cm(ts, 0), ntypes);
        }
        drop_device_copy();
    }

    anl_series *device() {
        if (!dev_) anl_check(anl_series_create(table_.data(), n_timestep(), (uint32_t)width_, &dev_));
        return dev_;
    }

   private:
    // Reads lines until the first row candidate (a trailing '\r' removed from
    // each); *header is the line before it.
    static void scan_to_first_row(std::ifstream &in, std::string *header, std::string *row) {
        std::string cur;
        while (in.good()) {
            *header = cur;
            std::getline(in, cur);
            if (cur.length() > 1 && cur.back() == '\r') cur.pop_back();
            if (logtext::is_row_candidate(cur)) break;
        }
        *row = cur;
    }

    // Parses the first candidate and every later line, in blocks of
    // nthreads * nbatch lines; thread k parses lines [k nbatch, (k+1) nbatch)
    // of the block into its own slots, then the accepted rows are appended
    // in order.
    void read_rows(std::ifstream &in, const std::string &first, unsigned int nthreads, unsigned int nbatch) {
        const size_t block = (size_t)nthreads * nbatch;
        std::vector<std::string> lines;
        lines.reserve(block);
        lines.push_back(first);
        std::vector<TFLOAT> parsed(block * width_, (TFLOAT)0);
        std::vector<unsigned int> steps(block, 0);
        std::vector<char> ok(block, 0);
        size_t accepted = 0;
        auto parse_range = [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; i++)
                ok[i] = logtext::parse_row(lines[i], cols_, &parsed[i * width_], &steps[i]) ? 1 : 0;
        };
        for (;;) {
            std::string s;
            while (lines.size() < block && in.good()) {
                std::getline(in, s);
                lines.push_back(s);
            }
            if (lines.empty()) break;
            if (nthreads == 1 || lines.size() < 2 * (size_t)nbatch) {
                parse_range(0, lines.size());
            } else {
                std::vector<std::thread> pool;
                for (unsigned int k = 0; k < nthreads; k++) {
                    const size_t lo = (size_t)k * nbatch, hi = std::min(lines.size(), lo + nbatch);
                    if (lo < hi) pool.emplace_back(parse_range, lo, hi);
                }
                for (auto &th : pool) th.join();
            }
            for (size_t i = 0; i < lines.size(); i++) {
                if (!ok[i]) continue;
                accepted++;  // 1-based count of accepted rows
                if ((accepted - 1) % keep_every_ != 0) continue;
                table_.insert(table_.end(), parsed.begin() + i * width_, parsed.begin() + (i + 1) * width_);
                step_of_row_.push_back(steps[i] > 0 ? steps[i] : (unsigned int)accepted);
            }
            lines.clear();
            if (!in.good()) break;
        }
    }

    size_t current_index(const std::string &spec) const {
        for (size_t k = 0; k < currents_.size(); k++)
            if (currents_[k].spec == spec) return k;
        std::cerr << "Error: cannot find the header \"" << spec << "\" in the list '";
        for (const auto &c : currents_) std::cerr << " '" << c.spec << "'";
        std::cerr << "\n";
        throw std::runtime_error("Wrong header name");
    }

    // J_k = sum over types (ascending) of v_cm(type) q_k(type), per component
    void fill_currents_of_row(unsigned int ts, const double *v_cm, size_t ntypes) {
        TFLOAT *row = &table_[(size_t)ts * width_ + cols_.stored()];
        for (size_t k = 0; k < currents_.size(); k++) {
            TFLOAT *J = row + 3 * k;
            J[0] = J[1] = J[2] = 0.0;
            for (size_t type = 0; type < ntypes; type++)
                for (int c = 0; c < 3; c++) J[c] += v_cm[type * 3 + c] * currents_[k].charge[type];
        }
    }

    void drop_device_copy() {  // the table changed: the device copy is rebuilt on next use
        if (dev_) {
            anl_series_destroy(dev_);
            dev_ = nullptr;
        }
    }

    Trajectory *traj_;
    unsigned int keep_every_;
    logtext::ColumnMap cols_;
    std::vector<logtext::ChargeCurrent<TFLOAT>> currents_;
    size_t width_ = 0;                     // stored columns + 3 per current
    std::vector<TFLOAT> table_;            // [rows][width_]
    std::vector<unsigned int> step_of_row_;
    anl_series *dev_ = nullptr;
};

}  // namespace analisi
