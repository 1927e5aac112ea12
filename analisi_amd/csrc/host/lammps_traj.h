// Trajectory: the memory-mapped LAMMPS-binary trajectory of rikigigi/analisi
// v0.5.2 (trajectory.{h,cpp}, lammps_struct.h), reading a window of
// timesteps into host arrays in atom order; device() mirrors the loaded
// window on the GPU for Gofrt / MSD.  Same constructor, window calls
// (set_data_access_block_size, set_access_at), accessors with absolute
// timesteps, error messages and on-disk format as the reference.
#pragma once

#include <fcntl.h>
#include <future>
#include <thread>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "analisi.h"

namespace analisi {

namespace lammps_bin {

// sizeof(Intestazione_timestep) of the reference on x86-64, with its padding:
// the end-of-file guards (lammps_struct.h:199-201, trajectory.cpp:250-252) use it
constexpr size_t HEADER_GUARD = 104;

template <class T>
const char *take(const char *p, const char *end, T *dst, size_t n = 1) {
    if (p + sizeof(T) * n > end) throw std::runtime_error("Error: end of file reached");
    std::memcpy(dst, p, sizeof(T) * n);
    return p + sizeof(T) * n;
}

// One timestep header in any of the three layouts (lammps_struct.h:31-62
// old, :65-97 old triclinic, :128-187 the 2020 revision with a magic string).
struct Header {
    int64_t timestep = 0, natoms = 0;
    int triclinic = 0;
    int boundary[6] = {0, 0, 0, 0, 0, 0};
    double box[6] = {0, 0, 0, 0, 0, 0};  // xlo, xhi, ylo, yhi, zlo, zhi
    double tilt[3] = {0, 0, 0};          // xy, xz, yz
    int size_one = 0, nchunk = 0;

    // parses at begin; returns the header's size in bytes (lammps_struct.h:196-216)
    size_t read(const char *begin, const char *end) {
        if (begin + HEADER_GUARD >= end) throw std::runtime_error("Error: end of file reached");
        const char *p = begin;
        int64_t first;
        take(p, end, &first);
        if (first < 0) {
            p += sizeof(int64_t) + (size_t)(-first);  // magic string
            int endian, revision;
            p = take(p, end, &endian);
            p = take(p, end, &revision);
            p = take(p, end, &timestep);
            p = take(p, end, &natoms);
            p = take(p, end, &triclinic);
            p = take(p, end, boundary, 6);
            p = take(p, end, box, 6);
            if (triclinic) p = take(p, end, tilt, 3);
            p = take(p, end, &size_one);
            if (revision > 0x0001) {
                int unit_len;
                p = take(p, end, &unit_len);
                if (unit_len > 0) p += unit_len;
                char time_flag;
                p = take(p, end, &time_flag);
                if (time_flag) p += sizeof(double);
                int columns_len;
                p = take(p, end, &columns_len);
                p += columns_len;
            }
            p = take(p, end, &nchunk);
        } else {
            p = take(p, end, &timestep);
            p = take(p, end, &natoms);
            p = take(p, end, &triclinic);
            p = take(p, end, boundary, 6);
            p = take(p, end, box, 6);
            if (triclinic) p = take(p, end, tilt, 3);
            p = take(p, end, &size_one);
            p = take(p, end, &nchunk);
        }
        if (p >= end) throw std::runtime_error("Error: end of file reached");
        if (size_one != NDOUBLE_ATOMO)
            throw std::runtime_error("ERROR: the binary format does not have " + std::to_string(NDOUBLE_ATOMO) +
                                     " numbers per atom but it has " + std::to_string(size_one));
        return (size_t)(p - begin);
    }
};

}  // namespace lammps_bin

class Trajectory {
   public:
    enum Errori { non_inizializzato = 0, oltre_fine_file = 2, Ok = 1 };

    // trajectory.cpp:42-127: map the file, read the first timestep's atoms
    // (ids in order of appearance, types), estimate the number of timesteps
    // from the file size over the first timestep's size.
    explicit Trajectory(const std::string &filename) {
        fd = open(filename.c_str(), O_RDONLY);
        if (fd == -1) throw std::runtime_error("Error opening the trajectory \"" + filename + "\"\n");
        struct stat sb;
        if (fstat(fd, &sb) == -1) {
            close(fd);
            fd = -1;
            throw std::runtime_error("Error in finding trajectory file size \"" + filename + "\"\n");
        }
        fsize = (size_t)sb.st_size;
        if (fsize == 0) {
            close(fd);
            fd = -1;
            throw std::runtime_error("mmap failed for file \"" + filename + "\".\n");
        }
        void *m = mmap(nullptr, fsize, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            close(fd);
            fd = -1;
            throw std::runtime_error("mmap failed for file \"" + filename + "\".\n");
        }
        file = (const char *)m;
        try {
            lammps_bin::Header h0;
            const size_t size0 = read_step(0, h0, nullptr);
            if (h0.natoms <= 0 || h0.natoms > (int64_t)INT32_MAX)
                throw std::runtime_error("Error: invalid number of atoms in the first timestep");
            natoms = (size_t)h0.natoms;
            triclinic = h0.triclinic != 0;  // the first timestep's cell shape holds for the file
            box_stride = triclinic ? 9 : 6;
            init_types();
            n_timesteps = fsize / size0;
            offsets.assign(1, 0);
            ok = true;
        } catch (...) {
            munmap((void *)file, fsize);
            close(fd);
            throw;
        }
    }
    ~Trajectory() {
        cancel_prefetch();
        if (dev) anl_traj_destroy(dev);
        if (file) munmap((void *)file, fsize);
        if (fd != -1) close(fd);
    }
    Trajectory(const Trajectory &) = delete;
    Trajectory &operator=(const Trajectory &) = delete;

    // trajectory.cpp:348-378: the window's size in timesteps
    Errori set_data_access_block_size(size_t n) {
        if (!ok) return non_inizializzato;
        if (loaded_timesteps == n) return Ok;
        cancel_prefetch();
        dati_caricati = false;
        loaded_timesteps = n;
        pos_.assign(n * natoms * 3, 0.0);
        vel_.assign(n * natoms * 3, 0.0);
        pcm_.assign(n * ntypes * 3, 0.0);
        vcm_.assign(n * ntypes * 3, 0.0);
        boxes.assign(n * box_stride, 0.0);
        return Ok;
    }

    // trajectory.cpp:458-650: reads timesteps [ts, ts + window) into the
    // window: boxes to the internal form, positions and velocities by atom id,
    // per-type centre of mass (running mean of the raw values, in file order),
    // then the optional pbc wrap.  A window read ahead by prefetch(ts) is
    // adopted instead of reading it again.
    Errori set_access_at(size_t ts) {
        if (!ok) return non_inizializzato;
        if (ts == current_timestep && dati_caricati) return Ok;
        if (ahead.valid()) {
            Window w = ahead.get();  // rethrows the reader's error
            if (w.start == ts && w.pos.size() == loaded_timesteps * natoms * 3) {
                adopt(w);
                return Ok;
            }
            if (w.dev) anl_traj_destroy(w.dev);
        }
        dati_caricati = false;
        Window w;
        w.start = ts;
        load_window(window_offsets(ts), w);
        adopt(w);
        return Ok;
    }

    // Starts reading the window at ts, and building its device mirror, on a
    // host thread: the reads, the host-side ordering and the uploads (on a
    // stream of their own) overlap the work on the current window.  The next
    // set_access_at(ts) adopts it; another prefetch or a window size change
    // discards it.  A hint only: results never depend on it.
    void prefetch(size_t ts) {
        if (!ok || loaded_timesteps == 0 || !prefetching) return;
        if (ts + loaded_timesteps > n_timesteps) return;
        if (ahead.valid() && ahead_start == ts) return;
        cancel_prefetch();
        ahead_start = ts;
        std::vector<size_t> offs = window_offsets(ts);  // indexes the file here, not on the reader
        ahead = std::async(std::launch::async, [this, ts, offs]() {
            Window w;
            w.start = ts;
            load_window(offs, w);
            std::vector<int32_t> tid(type_id.begin(), type_id.end());
            anl_check(anl_traj_create(w.pos.data(), w.boxes.data(), (uint32_t)box_stride, tid.data(),
                                      (uint32_t)ntypes, w.pcm.data(), loaded_timesteps, natoms, &w.dev));
            return w;
        });
    }
    void cancel_prefetch() {
        if (!ahead.valid()) return;
        try {
            Window w = ahead.get();
            if (w.dev) anl_traj_destroy(w.dev);
        } catch (...) {
            // a discarded window's error does not concern the caller
        }
    }
    // prefetch() on or off (on by default)
    void set_prefetch(bool on) {
        prefetching = on;
        if (!on) cancel_prefetch();
    }

    void set_pbc_wrap(bool p) {
        cancel_prefetch();
        wrap_pbc = p;
    }
    bool get_pbc_wrap() const { return wrap_pbc; }

    // trajectory.h:44-104: absolute timesteps; a timestep outside the window
    // moves the window there first
    double *positions(size_t t, size_t i) { return at(pos_.data(), t, i, 3 * natoms, 3, true); }
    double *velocity(size_t t, size_t i) { return at(vel_.data(), t, i, 3 * natoms, 3, true); }
    double *box(size_t t) { return at(boxes.data(), t, 0, box_stride, 0, false); }
    double *positions_cm(size_t t, size_t type) { return at(pcm_.data(), t, type, 3 * ntypes, 3, false); }
    double *velocity_cm(size_t t, size_t type) { return at(vcm_.data(), t, type, 3 * ntypes, 3, false); }

    unsigned int get_type(unsigned int i) const {
        if (i < natoms) return type_id[i];
        throw std::runtime_error("Atom index out of range\n");
    }
    std::vector<unsigned int> get_types() const { return types; }
    size_t get_ntypes() const { return ntypes; }
    size_t get_natoms() const { return natoms; }
    size_t get_ntimesteps() const { return n_timesteps; }
    size_t get_nloaded_timesteps() const { return loaded_timesteps; }
    // a window is sized and read (not only sized)
    bool has_window_data() const { return dati_caricati && loaded_timesteps > 0; }
    ssize_t get_current_timestep() const { return (ssize_t)current_timestep; }
    bool is_triclinic() const { return triclinic; }
    size_t get_box_stride() const { return box_stride; }
    // trajectory.cpp:204-226
    std::vector<int> get_lammps_id() const {
        std::vector<int> r(natoms, -1);
        for (auto &kv : id_map) r[kv.second] = kv.first;
        return r;
    }
    std::vector<int> get_lammps_type() const { return raw_types; }
    // trajectory.cpp:692-711: the LAMMPS timestep number of timestep t
    int64_t get_timestep_lammps(size_t t) {
        offset_of(t);
        return lammps_steps.at(t);
    }

    // basetrajectory.h:270-290: the buffer protocol serves positions or velocities
    void toggle_pos_vel() { serve_pos = !serve_pos; }
    bool serving_pos() const { return serve_pos; }
    double *positions_data() { return pos_.data(); }
    double *velocity_data() { return vel_.data(); }
    std::vector<ssize_t> get_shape() const { return {(ssize_t)loaded_timesteps, (ssize_t)natoms, 3}; }
    std::vector<ssize_t> get_stride() const {
        return {(ssize_t)(natoms * 3 * sizeof(double)), (ssize_t)(3 * sizeof(double)), (ssize_t)sizeof(double)};
    }

    double d2_minImage(size_t i, size_t j, size_t ti, size_t tj, double *x) {
        const double *c = box(ti);
        const double *xi = positions(ti, i), *xj = positions(tj, j);
        for (int k = 0; k < 3; k++) x[k] = xi[k] - xj[k];
        Trajectory_numpy::minImage_triclinic(x, c + 3, c + 6, triclinic);
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) d2 += x[k] * x[k];
        return d2;
    }

    void dump_lammps_bin_traj(const std::string &fname, int start_ts, int stop_ts) {
        lammps_bin::dump(*this, fname, start_ts, stop_ts);
    }

    // Window offset of a calculation on timesteps [primo, primo + need): the
    // device mirror holds the loaded window only, so the window is moved to
    // cover the range when it does not and is large enough.
    size_t window_offset(size_t primo, size_t need) {
        if (!(dati_caricati && primo >= current_timestep && primo + need <= current_timestep + loaded_timesteps)) {
            if (need > loaded_timesteps) {
                std::stringstream ss;
                ss << "there are not enough loaded timesteps inside the trajectory object. I need at least " << need
                   << " timesteps to do the requested calculation";
                throw std::runtime_error(ss.str());
            }
            // the window starts at primo, or earlier when it would run past
            // the last timestep
            size_t start = primo;
            if (n_timesteps >= loaded_timesteps) start = std::min(primo, n_timesteps - loaded_timesteps);
            if (set_access_at(start) != Ok) throw std::runtime_error("Error loading the file\n");
        }
        if (next_set) {  // the caller's next calculation: read its window ahead
            next_set = false;
            if (!(dati_caricati && next_start == current_timestep)) prefetch(next_start);
        }
        return primo - current_timestep;
    }
    // The first timestep of the calculation after the current one (block
    // loops know it): once the current calculation's window is in place, the
    // window that one will need is read ahead with prefetch().
    void set_next_window(size_t primo) {
        if (loaded_timesteps == 0 || loaded_timesteps > n_timesteps) return;
        next_start = std::min(primo, n_timesteps - loaded_timesteps);  // window_offset's rule
        next_set = true;
    }
    // changes whenever the window's contents change; operators rebuild their
    // device handles when it does
    uint64_t generation() const { return generation_; }

    // device mirror of the loaded window, rebuilt after the window moves
    anl_traj *device() {
        if (!dati_caricati || loaded_timesteps == 0)
            throw std::runtime_error("no timesteps are loaded: call setAccessWindowSize and setAccessStart first");
        if (!dev || dev_generation != generation_) {
            if (dev) anl_traj_destroy(dev);
            dev = nullptr;
            std::vector<int32_t> tid(type_id.begin(), type_id.end());
            anl_check(anl_traj_create(pos_.data(), boxes.data(), (uint32_t)box_stride, tid.data(), (uint32_t)ntypes,
                                      pcm_.data(), loaded_timesteps, natoms, &dev));
            dev_generation = generation_;
        }
        return dev;
    }

   private:
    // one window's host arrays (and, when read ahead, its device mirror)
    struct Window {
        size_t start = 0;
        std::vector<double> pos, vel, pcm, vcm, boxes;
        anl_traj *dev = nullptr;
    };
    std::vector<size_t> window_offsets(size_t ts) {
        std::vector<size_t> offs(loaded_timesteps);
        for (size_t k = 0; k < loaded_timesteps; k++) offs[k] = offset_of(ts + k);
        return offs;
    }
    // reads the timesteps at offs into w; touches no member (safe on the
    // prefetch thread: the file, the atom map and the sizes are read-only
    // once the window size is set)
    void load_window(const std::vector<size_t> &offs, Window &w) const {
        const size_t nl = offs.size();
        w.pos.assign(nl * natoms * 3, 0.0);
        w.vel.assign(nl * natoms * 3, 0.0);
        w.pcm.assign(nl * ntypes * 3, 0.0);
        w.vcm.assign(nl * ntypes * 3, 0.0);
        w.boxes.assign(nl * box_stride, 0.0);
        // timesteps are independent (their own slices and centre-of-mass
        // counts): split them over up to 8 host threads
        const size_t nth = std::min<size_t>(nl, std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
        if (nth <= 1) {
            load_steps(offs, w, 0, nl);
            return;
        }
        std::vector<std::future<void>> parts;
        for (size_t p = 0; p < nth; p++)
            parts.push_back(std::async(std::launch::async, [this, &offs, &w, p, nth, nl] {
                load_steps(offs, w, nl * p / nth, nl * (p + 1) / nth);
            }));
        for (auto &f : parts) f.get();  // rethrows a step's error (the other parts are joined on the way out)
    }
    void load_steps(const std::vector<size_t> &offs, Window &w, size_t k0, size_t k1) const {
        std::vector<size_t> cont(ntypes);
        for (size_t k = k0; k < k1; k++) {
            lammps_bin::Header h;
            std::vector<std::pair<const char *, int>> chunks;
            read_step(offs[k], h, &chunks);
            double *c = &w.boxes[k * box_stride];
            for (int i = 0; i < 6; i++) c[i] = h.box[i];
            if (triclinic)
                for (int i = 0; i < 3; i++) c[6 + i] = h.tilt[i];
            Trajectory_numpy::lammps_to_internal(c);
            double *pc = &w.pcm[k * ntypes * 3], *vc = &w.vcm[k * ntypes * 3];
            std::fill(cont.begin(), cont.end(), 0);
            for (auto &ch : chunks)
                for (int ia = 0; ia < ch.second; ia++) {
                    double row[lammps_bin::NDOUBLE_ATOMO];
                    std::memcpy(row, ch.first + (size_t)ia * sizeof(row), sizeof(row));
                    auto it = id_map.find((int)std::round(row[0]));
                    if (it == id_map.end()) throw std::out_of_range("map::at");
                    const size_t id = it->second;
                    double *x = &w.pos[(k * natoms + id) * 3], *v = &w.vel[(k * natoms + id) * 3];
                    for (int q = 0; q < 3; q++) {
                        x[q] = row[2 + q];
                        v[q] = row[5 + q];
                    }
                    const unsigned int t = type_id[id];
                    cont[t]++;
                    for (int q = 0; q < 3; q++) {
                        pc[3 * t + q] += (x[q] - pc[3 * t + q]) / (double)cont[t];
                        vc[3 * t + q] += (v[q] - vc[3 * t + q]) / (double)cont[t];
                    }
                }
            if (wrap_pbc) pbc_wrap(&w.pos[k * natoms * 3], c);
        }
    }
    void adopt(Window &w) {
        pos_.swap(w.pos);
        vel_.swap(w.vel);
        pcm_.swap(w.pcm);
        vcm_.swap(w.vcm);
        boxes.swap(w.boxes);
        dati_caricati = true;
        current_timestep = w.start;
        generation_++;
        if (w.dev) {
            if (dev) anl_traj_destroy(dev);
            dev = w.dev;
            dev_generation = generation_;
        }
    }

    // trajectory.cpp:237-290: header and chunk table of the timestep at off;
    // returns its size in bytes
    size_t read_step(size_t off, lammps_bin::Header &h, std::vector<std::pair<const char *, int>> *chunks) const {
        if (off + lammps_bin::HEADER_GUARD > fsize) throw std::runtime_error("Error: trying to read beyond the end of file");
        const char *end = file + fsize;
        size_t n = h.read(file + off, end);
        if (natoms != 0 && (size_t)h.natoms != natoms) {
            std::stringstream ss;
            ss << "Error: at the LAMMPS timestep " << h.timestep << " the number of atoms has changed \n";
            throw std::runtime_error(ss.str());
        }
        if (h.nchunk <= 0) throw std::runtime_error("Error: the number of chunks of the timestep cannot be <= 0");
        for (int i = 0; i < h.nchunk; i++) {
            int ndouble;
            lammps_bin::take(file + off + n, end, &ndouble);
            if (ndouble % lammps_bin::NDOUBLE_ATOMO != 0)
                throw std::runtime_error("Number of bytes of the chunk is not a multiple of the number of atoms\n");
            n += sizeof(int);
            const int na = ndouble / lammps_bin::NDOUBLE_ATOMO;
            if (na < 0 || off + n + (size_t)na * sizeof(double) * lammps_bin::NDOUBLE_ATOMO > fsize)
                throw std::runtime_error("Error: trying to read beyond the end of file");
            if (chunks) chunks->push_back({file + off + n, na});
            n += (size_t)na * sizeof(double) * lammps_bin::NDOUBLE_ATOMO;
        }
        return n;
    }
    // trajectory.cpp:130-201: atom k is the k-th distinct id of the first
    // timestep (a repeated id is skipped, as there); types by id
    void init_types() {
        lammps_bin::Header h;
        std::vector<std::pair<const char *, int>> chunks;
        read_step(0, h, &chunks);
        id_map.clear();
        int next = 0;
        raw_types.assign(natoms, -1);
        for (auto &ch : chunks)
            for (int ia = 0; ia < ch.second; ia++) {
                if ((size_t)next >= natoms)
                    throw std::runtime_error(
                        "Error: the number of atoms does not correspond to the sum of the number of atoms of each "
                        "chunk\n");
                double row[2];
                std::memcpy(row, ch.first + (size_t)ia * sizeof(double) * lammps_bin::NDOUBLE_ATOMO, sizeof(row));
                const int id = (int)std::round(row[0]);
                if (id < 0) throw std::runtime_error("Error: found a negative atomic id in the trajectory");
                if (id_map.find(id) == id_map.end()) id_map[id] = next++;
            }
        for (auto &ch : chunks)
            for (int ia = 0; ia < ch.second; ia++) {
                double row[2];
                std::memcpy(row, ch.first + (size_t)ia * sizeof(double) * lammps_bin::NDOUBLE_ATOMO, sizeof(row));
                raw_types[id_map.at((int)std::round(row[0]))] = (int)std::round(row[1]);
            }
        // basetrajectory.cpp:51-89: type index = rank among the distinct types
        std::vector<int> distinct;
        for (int t : raw_types)
            if (std::find(distinct.begin(), distinct.end(), t) == distinct.end()) distinct.push_back(t);
        std::sort(distinct.begin(), distinct.end());
        types.assign(distinct.begin(), distinct.end());
        ntypes = distinct.size();
        type_id.resize(natoms);
        for (size_t i = 0; i < natoms; i++)
            type_id[i] = (unsigned int)(std::lower_bound(distinct.begin(), distinct.end(), raw_types[i]) -
                                        distinct.begin());
    }
    // byte offset of timestep t, indexing the timesteps before it on demand
    // (trajectory.cpp:494-505)
    size_t offset_of(size_t t) {
        while (offsets.size() <= t) {
            const size_t last = offsets.size() - 1;
            lammps_bin::Header h;
            const size_t n = read_step(offsets[last], h, nullptr);
            if (lammps_steps.size() <= last) lammps_steps.resize(last + 1);
            lammps_steps[last] = h.timestep;
            offsets.push_back(offsets[last] + n);
        }
        if (lammps_steps.size() <= t) {
            lammps_bin::Header h;
            read_step(offsets[t], h, nullptr);
            lammps_steps.resize(t + 1);
            lammps_steps[t] = h.timestep;
        }
        return offsets[t];
    }
    double *at(double *base, size_t t, size_t i, size_t s1, size_t s2, bool atom) {
        if (atom && i >= natoms) {
            std::stringstream ss;
            ss << "Requested atom index (" << i << ") is not in the range [0, " << natoms - 1 << "]\n";
            throw std::runtime_error(ss.str());
        }
        if (dati_caricati && t >= current_timestep && t - current_timestep < loaded_timesteps)
            return base + (t - current_timestep) * s1 + i * s2;
        if (set_access_at(t) != Ok) throw std::runtime_error("Error loading the file\n");
        if (t >= current_timestep && t - current_timestep < loaded_timesteps)
            return base + (t - current_timestep) * s1 + i * s2;
        throw std::runtime_error("requested timestep is out of range");
    }
    // basetrajectory.h:145-161: one frame's positions p around the cell c
    void pbc_wrap(double *p, const double *c) const {
        const double tilt0[3] = {0, 0, 0};
        for (size_t ia = 0; ia < natoms; ++ia) {
            double *xa = &p[ia * 3];
            for (int q = 0; q < 3; ++q) xa[q] = xa[q] - c[3 + q];
            Trajectory_numpy::minImage_triclinic(xa, c + 3, triclinic ? c + 6 : tilt0, triclinic);
            for (int q = 0; q < 3; ++q) xa[q] = xa[q] + c[3 + q];
        }
    }

    int fd = -1;
    const char *file = nullptr;
    size_t fsize = 0;
    bool ok = false, dati_caricati = false, triclinic = false, wrap_pbc = false, serve_pos = true;
    size_t natoms = 0, ntypes = 0, n_timesteps = 0, loaded_timesteps = 0, current_timestep = 0, box_stride = 6;
    std::vector<size_t> offsets;
    std::vector<int64_t> lammps_steps;
    std::map<int, int> id_map;
    std::vector<int> raw_types;
    std::vector<unsigned int> type_id, types;
    std::vector<double> pos_, vel_, pcm_, vcm_, boxes;
    uint64_t generation_ = 0, dev_generation = 0;
    anl_traj *dev = nullptr;
    std::future<Window> ahead;  // window being read ahead by prefetch()
    size_t ahead_start = 0, next_start = 0;
    bool prefetching = true, next_set = false;
};

}  // namespace analisi
