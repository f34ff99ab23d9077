// state_io.hip -- the state's files (Source/StateIO.cpp): AMReX plotfiles
// (HyperCLaw-V1.1 header, VisMF Cell_H and FAB data) readable by AMReX and
// its tools, and checkpoints in the reference's own format (a text Header
// with the mesh in hex floats and the BoxArray, then every field as a VisMF
// MultiFab with its ghosts), so a run here and a GUST run on AMReX can
// exchange fields and a served request can restart bit for bit.
//
// Host code: every field is copied to the host once (one synchronisation),
// then written box by box; FAB data is little-endian IEEE doubles,
// component-major, i fastest, as AMReX's NativeFormat writes on x86 and
// gfx950 hosts.
#include "common.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace gust_amd {

namespace {

namespace fs = std::filesystem;

constexpr const char* kCheckpointMagic = "gust-checkpoint";
constexpr int kCheckpointVersion = 1;
constexpr const char* kFabDescriptor = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

const int kLocs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};
const char* kNames[6] = {"u", "v", "w", "p", "theta", "tke"};

// A field on the host: its view's box [lo, hi] and the values (i fastest).
struct HostField {
    IBox box;
    std::vector<double> v;
    size_t idx(int i, int j, int k) const
    {
        return size_t(i - box.lo[0]) + size_t(box.len(0)) * (size_t(j - box.lo[1]) + size_t(box.len(1)) *
                                                                                     size_t(k - box.lo[2]));
    }
    double at(int i, int j, int k) const { return v[idx(i, j, k)]; }
    double& at(int i, int j, int k) { return v[idx(i, j, k)]; }
};

IBox view_box(const gust_array4& a)
{
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = a.begin[d];
        b.hi[d] = a.end[d] - 1;
    }
    return b;
}

// The view's component 0 on the host (any strides).
HostField download(const gust_array4& a, hipStream_t st)
{
    HostField h;
    h.box = view_box(a);
    h.v.resize(size_t(h.box.npts()));
    const size_t row = size_t(h.box.len(0)) * sizeof(double);
    for (int k = h.box.lo[2]; k <= h.box.hi[2]; ++k) {
        GUST_HIP(hipMemcpy2DAsync(&h.at(h.box.lo[0], h.box.lo[1], k), row,
                                  static_cast<const double*>(a.p) + offset_of(a, h.box.lo[0], h.box.lo[1], k),
                                  size_t(a.jstride) * sizeof(double), row, size_t(h.box.len(1)),
                                  hipMemcpyDeviceToHost, st));
    }
    return h;
}

void upload(const HostField& h, const gust_array4& a, hipStream_t st)
{
    const size_t row = size_t(h.box.len(0)) * sizeof(double);
    for (int k = h.box.lo[2]; k <= h.box.hi[2]; ++k) {
        GUST_HIP(hipMemcpy2DAsync(static_cast<double*>(a.p) + offset_of(a, h.box.lo[0], h.box.lo[1], k),
                                  size_t(a.jstride) * sizeof(double), h.v.data() + h.idx(h.box.lo[0], h.box.lo[1], k), row, row,
                                  size_t(h.box.len(1)), hipMemcpyHostToDevice, st));
    }
}

std::string box_text(const IBox& b, int loc)
{
    char s[160];
    std::snprintf(s, sizeof s, "((%d,%d,%d) (%d,%d,%d) (%d,%d,%d))", b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1],
                  b.hi[2], loc & 1, (loc >> 1) & 1, (loc >> 2) & 1);
    return s;
}

// AMReX's BoxList::maxSize along x and y: ceil(n / m) pieces, the first
// n mod pieces one cell longer; every box spans z.
std::vector<IBox> chop(const gust_geometry& g, int max_grid_size)
{
    std::vector<std::pair<int, int>> px, py;
    for (int d = 0; d < 2; ++d) {
        const int n = g.hi[d] - g.lo[d] + 1;
        const int pieces = max_grid_size > 0 ? (n + max_grid_size - 1) / max_grid_size : 1;
        const int size = n / pieces, extra = n % pieces;
        int at = g.lo[d];
        auto& out = d == 0 ? px : py;
        for (int p = 0; p < pieces; ++p) {
            const int len = size + (p < extra ? 1 : 0);
            out.push_back({at, at + len - 1});
            at += len;
        }
    }
    std::vector<IBox> boxes;
    for (const auto& y : py) {
        for (const auto& x : px) { boxes.push_back(IBox{{x.first, y.first, g.lo[2]}, {x.second, y.second, g.hi[2]}}); }
    }
    return boxes;
}

IBox at_loc(IBox b, int loc, const gust_geometry& g)
{
    // A box's faces of type loc: the high face is the box's own where the
    // box ends at the domain's high side or loc is nodal there (AMReX's
    // convert, surroundingNodes: every box owns both its faces).
    for (int d = 0; d < 3; ++d) {
        if ((loc >> d) & 1) { b.hi[d] += 1; }
    }
    (void)g;
    return b;
}

void make_dirs(const std::string& dir)
{
    std::error_code ec;
    fs::create_directories(dir, ec);
    if (ec) { GUST_FAIL("gust: could not create the directory %s: %s", dir.c_str(), ec.message().c_str()); }
}

std::ofstream open_out(const std::string& path, std::ios::openmode mode = std::ios::out)
{
    std::ofstream os(path, mode);
    if (!os.good()) { GUST_FAIL("gust: could not open %s for writing", path.c_str()); }
    return os;
}

std::string g17(double x)
{
    char s[64];
    std::snprintf(s, sizeof s, "%.17g", x);
    return s;
}

std::string e17(double x)
{
    char s[64];
    std::snprintf(s, sizeof s, "%.17e", x);
    return s;
}

std::string hexf(double x)
{
    char s[64];
    std::snprintf(s, sizeof s, "%a", x);
    return s;
}

// A VisMF MultiFab (header version 1, one data file): comps[n] sampled on
// each box grown by ngrow (FAB data), min and max over each valid box.
void write_vismf(const std::string& base, const std::vector<IBox>& boxes, int loc, int ngrow,
                 const std::vector<const HostField*>& comps)
{
    const fs::path p(base);
    const std::string dfile = p.filename().string() + "_D_00000";
    std::ofstream data = open_out(base + "_D_00000", std::ios::out | std::ios::binary);
    std::vector<long long> offsets;
    const int nc = int(comps.size());
    std::vector<std::vector<double>> mins, maxs;
    for (const IBox& b : boxes) {
        const IBox g = grow(b, ngrow, ngrow, ngrow);
        offsets.push_back(static_cast<long long>(data.tellp()));
        data << kFabDescriptor << box_text(g, loc) << " " << nc << "\n";
        std::vector<double> mn(size_t(nc), 0.0), mx(size_t(nc), 0.0);
        std::vector<double> buf(size_t(g.npts()));
        for (int c = 0; c < nc; ++c) {
            size_t at = 0;
            bool first = true;
            for (int k = g.lo[2]; k <= g.hi[2]; ++k)
                for (int j = g.lo[1]; j <= g.hi[1]; ++j)
                    for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
                        const double x = comps[size_t(c)]->at(i, j, k);
                        buf[at++] = x;
                        const bool valid = i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] &&
                                           k >= b.lo[2] && k <= b.hi[2];
                        if (valid) {
                            if (first || x < mn[size_t(c)]) { mn[size_t(c)] = x; }
                            if (first || x > mx[size_t(c)]) { mx[size_t(c)] = x; }
                            first = false;
                        }
                    }
            data.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size() * sizeof(double)));
        }
        mins.push_back(mn);
        maxs.push_back(mx);
    }
    if (!data.good()) { GUST_FAIL("gust: could not write %s_D_00000", base.c_str()); }
    std::ofstream h = open_out(base + "_H");
    h << "1\n1\n" << nc << "\n" << ngrow << "\n(" << boxes.size() << " 0\n";
    for (const IBox& b : boxes) { h << box_text(b, loc) << "\n"; }
    h << ")\n" << boxes.size() << "\n";
    for (size_t n = 0; n < boxes.size(); ++n) { h << "FabOnDisk: " << dfile << " " << offsets[n] << "\n"; }
    h << "\n" << boxes.size() << "," << nc << "\n";
    for (const auto& m : mins) {
        for (const double x : m) { h << e17(x) << ","; }
        h << "\n";
    }
    h << "\n" << boxes.size() << "," << nc << "\n";
    for (const auto& m : maxs) {
        for (const double x : m) { h << e17(x) << ","; }
        h << "\n";
    }
    h << "\n";
    if (!h.good()) { GUST_FAIL("gust: could not write %s_H", base.c_str()); }
}

// One FAB of a VisMF MultiFab back into h (the FAB's box clipped to h's).
void read_vismf(const std::string& base, int ncomp_expected, std::vector<HostField*> out, const char* what,
                const std::string& dir)
{
    std::vector<char> seen(out[0]->v.size(), 0);
    std::ifstream h(base + "_H");
    if (!h.good()) { GUST_FAIL("gust: io.restart = %s: missing %s_H", dir.c_str(), what); }
    std::string text((std::istreambuf_iterator<char>(h)), std::istreambuf_iterator<char>());
    size_t at = 0;
    int nfab = 0;
    while ((at = text.find("FabOnDisk: ", at)) != std::string::npos) {
        at += 11;
        const size_t sp = text.find(' ', at), nl = text.find('\n', at);
        if (sp == std::string::npos || nl == std::string::npos) { break; }
        const std::string file = text.substr(at, sp - at);
        const long long off = std::atoll(text.substr(sp + 1, nl - sp - 1).c_str());
        std::ifstream d(fs::path(base).parent_path() / file, std::ios::binary);
        if (!d.good()) { GUST_FAIL("gust: io.restart = %s: missing the data file %s", dir.c_str(), file.c_str()); }
        d.seekg(off);
        std::string line;
        std::getline(d, line);
        int b[9], nc = 0;
        const size_t pb = line.find(")))((");
        if (line.rfind(kFabDescriptor, 0) != 0 || pb == std::string::npos ||
            std::sscanf(line.c_str() + pb + 3, "((%d,%d,%d) (%d,%d,%d) (%d,%d,%d)) %d", &b[0], &b[1], &b[2], &b[3],
                        &b[4], &b[5], &b[6], &b[7], &b[8], &nc) != 10) {
            GUST_FAIL("gust: io.restart = %s: %s: not a FAB of 8-byte little-endian reals", dir.c_str(), what);
        }
        if (nc != ncomp_expected) {
            GUST_FAIL("gust: io.restart = %s: %s has %d components, expected %d", dir.c_str(), what, nc,
                      ncomp_expected);
        }
        const IBox fb{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
        std::vector<double> buf(size_t(fb.npts()));
        for (int c = 0; c < nc; ++c) {
            d.read(reinterpret_cast<char*>(buf.data()), std::streamsize(buf.size() * sizeof(double)));
            if (!d.good()) { GUST_FAIL("gust: io.restart = %s: %s: truncated FAB data", dir.c_str(), what); }
            HostField* dst = out[size_t(c)];
            size_t q = 0;
            for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
                for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
                    for (int i = fb.lo[0]; i <= fb.hi[0]; ++i, ++q) {
                        if (i >= dst->box.lo[0] && i <= dst->box.hi[0] && j >= dst->box.lo[1] &&
                            j <= dst->box.hi[1] && k >= dst->box.lo[2] && k <= dst->box.hi[2]) {
                            dst->at(i, j, k) = buf[q];
                            seen[dst->idx(i, j, k)] = 1;
                        }
                    }
        }
        ++nfab;
    }
    if (nfab == 0) { GUST_FAIL("gust: io.restart = %s: %s_H lists no FAB", dir.c_str(), what); }
    for (const char c : seen) {
        if (!c) { GUST_FAIL("gust: io.restart = %s: %s does not cover the domain", dir.c_str(), what); }
    }
}

std::string step_name(const char* prefix, int step)
{
    char s[32];
    std::snprintf(s, sizeof s, "%05d", step);
    return std::string(prefix) + s;
}

void put_name(const std::string& s, char* out, int64_t cap)
{
    if (!out || cap <= 0) { return; }
    const size_t n = std::min<size_t>(s.size(), size_t(cap - 1));
    std::memcpy(out, s.data(), n);
    out[n] = '\0';
}

void check_io_args(const char* op, const gust_geometry* g, const double* lo, const double* hi, const gust_state* q)
{
    if (!g || !lo || !hi || !q) { GUST_FAIL("gust: %s: null argument", op); }
    char name[32];
    for (int n = 0; n < 6; ++n) {
        std::snprintf(name, sizeof name, "state.%s", kNames[n]);
        check_field(op, name, q->f[n], *g, kLocs[n], 0);
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_write_plotfile(const char* prefix, int32_t step, const gust_geometry* geom, const double* prob_lo,
                        const double* prob_hi, const gust_state* q, const gust_array4* z_nd, const gust_array4* detJ,
                        int32_t max_grid_size, char* name, int64_t cap, void* stream)
{
    return guarded([&] {
        const char* op = "write_plotfile";
        check_io_args(op, geom, prob_lo, prob_hi, q);
        if (!prefix || !z_nd) { GUST_FAIL("gust: %s: null argument", op); }
        const gust_geometry& g = *geom;
        check_nodal_heights(op, z_nd, g);
        if (detJ) { check_field(op, "detJ", detJ, g, GUST_LOC_CELL, 0); }
        hipStream_t st = hipStream_t(stream);
        std::vector<HostField> f;
        for (int n = 0; n < 6; ++n) { f.push_back(download(*q->f[n], st)); }
        HostField z = download(*z_nd, st);
        HostField J;
        if (detJ) { J = download(*detJ, st); }
        GUST_HIP(hipStreamSynchronize(st));
        const std::string dir = step_name(prefix, step);
        make_dirs(dir + "/Level_0");
        // The cell variables (Source/StateIO.cpp:44-84): face means of u, v,
        // w, the cell fields, the mean of the eight node heights.
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        std::vector<HostField> plt(detJ ? 8 : 7);
        for (auto& p : plt) {
            p.box = cells;
            p.v.resize(size_t(cells.npts()));
        }
        for (int k = cells.lo[2]; k <= cells.hi[2]; ++k)
            for (int j = cells.lo[1]; j <= cells.hi[1]; ++j)
                for (int i = cells.lo[0]; i <= cells.hi[0]; ++i) {
                    const double half = 0.5, eighth = 0.125;
                    plt[0].at(i, j, k) = half * (f[0].at(i, j, k) + f[0].at(i + 1, j, k));
                    plt[1].at(i, j, k) = half * (f[1].at(i, j, k) + f[1].at(i, j + 1, k));
                    plt[2].at(i, j, k) = half * (f[2].at(i, j, k) + f[2].at(i, j, k + 1));
                    plt[3].at(i, j, k) = f[3].at(i, j, k);
                    plt[4].at(i, j, k) = f[4].at(i, j, k);
                    plt[5].at(i, j, k) = f[5].at(i, j, k);
                    plt[6].at(i, j, k) = eighth * (z.at(i, j, k) + z.at(i + 1, j, k) + z.at(i, j + 1, k) +
                                                   z.at(i + 1, j + 1, k) + z.at(i, j, k + 1) + z.at(i + 1, j, k + 1) +
                                                   z.at(i, j + 1, k + 1) + z.at(i + 1, j + 1, k + 1));
                    if (detJ) { plt[7].at(i, j, k) = J.at(i, j, k); }
                }
        const std::vector<IBox> boxes = chop(g, max_grid_size);
        std::vector<const HostField*> comps;
        for (const auto& p : plt) { comps.push_back(&p); }
        write_vismf(dir + "/Level_0/Cell", boxes, GUST_LOC_CELL, 0, comps);
        {
            std::ofstream h = open_out(dir + "/Header");
            const char* names[8] = {"u", "v", "w", "p", "theta", "tke", "z", "detJ"};
            h << "HyperCLaw-V1.1\n" << plt.size() << "\n";
            for (size_t n = 0; n < plt.size(); ++n) { h << names[n] << "\n"; }
            h << "3\n0\n0\n";
            for (int d = 0; d < 3; ++d) { h << g17(prob_lo[d]) << " "; }
            h << "\n";
            for (int d = 0; d < 3; ++d) { h << g17(prob_hi[d]) << " "; }
            h << "\n\n" << box_text(cells, GUST_LOC_CELL) << " \n" << step << " \n";
            for (int d = 0; d < 3; ++d) { h << g17(g.dx[d]) << " "; }
            h << "\n0\n0\n0 " << boxes.size() << " 0\n" << step << "\n";
            for (const IBox& b : boxes) {
                for (int d = 0; d < 3; ++d) {
                    h << g17(prob_lo[d] + g.dx[d] * b.lo[d]) << " " << g17(prob_lo[d] + g.dx[d] * (b.hi[d] + 1))
                      << "\n";
                }
            }
            h << "Level_0/Cell\n";
            if (!h.good()) { GUST_FAIL("gust: could not write %s/Header", dir.c_str()); }
        }
        // The node heights with their ghost nodes (StateIO.cpp:83).
        write_vismf(dir + "/Level_0/z_phys_nd", {at_loc(cells, 7, g)}, 7, 1, {&z});
        put_name(dir, name, cap);
    });
}

int gust_write_checkpoint(const char* prefix, int32_t step, const gust_geometry* geom, const double* prob_lo,
                          const double* prob_hi, const gust_state* q, const gust_array4* z_nd, int32_t max_grid_size,
                          char* name, int64_t cap, void* stream)
{
    return guarded([&] {
        const char* op = "write_checkpoint";
        check_io_args(op, geom, prob_lo, prob_hi, q);
        if (!prefix || !z_nd) { GUST_FAIL("gust: %s: null argument", op); }
        const gust_geometry& g = *geom;
        check_nodal_heights(op, z_nd, g);
        for (int n = 0; n < 6; ++n) {
            char w[32];
            std::snprintf(w, sizeof w, "state.%s", kNames[n]);
            check_field(op, w, q->f[n], g, kLocs[n], 2);
        }
        hipStream_t st = hipStream_t(stream);
        std::vector<HostField> f;
        for (int n = 0; n < 6; ++n) { f.push_back(download(*q->f[n], st)); }
        HostField z = download(*z_nd, st);
        GUST_HIP(hipStreamSynchronize(st));
        const std::string dir = step_name(prefix, step);
        make_dirs(dir + "/Level_0");
        const std::vector<IBox> boxes = chop(g, max_grid_size);
        {
            std::ofstream h = open_out(dir + "/Header");
            h << kCheckpointMagic << " " << kCheckpointVersion << "\n";
            h << "real_bytes 8\nstep " << step << "\n";
            h << "n_cell " << (g.hi[0] - g.lo[0] + 1) << " " << (g.hi[1] - g.lo[1] + 1) << " "
              << (g.hi[2] - g.lo[2] + 1) << "\n";
            h << "prob_lo";
            for (int d = 0; d < 3; ++d) { h << " " << hexf(prob_lo[d]); }
            h << "\nprob_hi";
            for (int d = 0; d < 3; ++d) { h << " " << hexf(prob_hi[d]); }
            h << "\nis_periodic " << g.periodic[0] << " " << g.periodic[1] << " " << g.periodic[2] << "\n";
            h << "(" << boxes.size() << " 0\n";
            for (const IBox& b : boxes) { h << box_text(b, GUST_LOC_CELL) << "\n"; }
            h << ")\n";
            if (!h.good()) { GUST_FAIL("gust: could not write %s/Header", dir.c_str()); }
        }
        for (int n = 0; n < 6; ++n) {
            std::vector<IBox> fb;
            for (const IBox& b : boxes) { fb.push_back(at_loc(b, kLocs[n], g)); }
            write_vismf(dir + "/Level_0/" + kNames[n], fb, kLocs[n], 2, {&f[size_t(n)]});
        }
        std::vector<IBox> nb;
        for (const IBox& b : boxes) { nb.push_back(at_loc(b, 7, g)); }
        write_vismf(dir + "/Level_0/z_phys_nd", nb, 7, 1, {&z});
        put_name(dir, name, cap);
    });
}

int gust_read_checkpoint(const char* dir, const gust_geometry* geom, const double* prob_lo, const double* prob_hi,
                         gust_state* q, gust_array4* z_nd, int32_t* step, void* stream)
{
    return guarded([&] {
        const char* op = "read_checkpoint";
        check_io_args(op, geom, prob_lo, prob_hi, q);
        if (!dir || !step) { GUST_FAIL("gust: %s: null argument", op); }
        const gust_geometry& g = *geom;
        if (z_nd) { check_nodal_heights(op, z_nd, g); }
        const std::string d(dir);
        std::ifstream h(d + "/Header");
        if (!h.good()) { GUST_FAIL("gust: io.restart = %s: no Header", dir); }
        std::string magic, key;
        int version = 0;
        h >> magic >> version;
        if (!h || magic != kCheckpointMagic || version != kCheckpointVersion) {
            GUST_FAIL("gust: io.restart = %s: not a GUST checkpoint of format version %d", dir, kCheckpointVersion);
        }
        auto expect = [&](const char* want) {
            h >> key;
            if (!h || key != want) { GUST_FAIL("gust: io.restart = %s: corrupt Header: expected '%s'", dir, want); }
        };
        auto real = [&](const char* what) {
            std::string tok;
            h >> tok;
            char* end = nullptr;
            const double x = std::strtod(tok.c_str(), &end);
            if (!h || tok.empty() || *end != '\0') {
                GUST_FAIL("gust: io.restart = %s: corrupt Header: bad value for '%s'", dir, what);
            }
            return x;
        };
        int rb = 0, s = 0, n[3] = {0, 0, 0}, per[3] = {0, 0, 0};
        expect("real_bytes");
        h >> rb;
        if (!h) { GUST_FAIL("gust: io.restart = %s: corrupt Header: bad value for 'real_bytes'", dir); }
        if (rb != 8) {
            GUST_FAIL("gust: io.restart = %s: checkpoint was written with %d-byte reals; this build uses 8", dir, rb);
        }
        expect("step");
        h >> s;
        if (!h) { GUST_FAIL("gust: io.restart = %s: corrupt Header: bad value for 'step'", dir); }
        expect("n_cell");
        h >> n[0] >> n[1] >> n[2];
        if (!h) { GUST_FAIL("gust: io.restart = %s: corrupt Header: bad value for 'n_cell'", dir); }
        double lo[3], hi[3];
        expect("prob_lo");
        for (double& x : lo) { x = real("prob_lo"); }
        expect("prob_hi");
        for (double& x : hi) { x = real("prob_hi"); }
        expect("is_periodic");
        h >> per[0] >> per[1] >> per[2];
        if (!h) { GUST_FAIL("gust: io.restart = %s: corrupt Header: bad value for 'is_periodic'", dir); }
        for (int k = 0; k < 3; ++k) {
            if (n[k] != g.hi[k] - g.lo[k] + 1) {
                GUST_FAIL("gust: io.restart = %s: checkpoint mesh.n_cell = (%d,%d,%d) differs from the inputs' "
                          "(%d,%d,%d)", dir, n[0], n[1], n[2], g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1,
                          g.hi[2] - g.lo[2] + 1);
            }
            if (lo[k] != prob_lo[k] || hi[k] != prob_hi[k]) {
                GUST_FAIL("gust: io.restart = %s: checkpoint mesh.prob_lo/prob_hi differ from the inputs", dir);
            }
            if (per[k] != g.periodic[k]) {
                GUST_FAIL("gust: io.restart = %s: checkpoint mesh.is_periodic differs from the inputs", dir);
            }
        }
        hipStream_t st = hipStream_t(stream);
        // Valid values exactly; of the ghosts only the node heights'.
        for (int k = 0; k < 6; ++k) {
            HostField f;
            f.box = loc_box(g, kLocs[k]);
            f.v.assign(size_t(f.box.npts()), std::nan(""));
            read_vismf(d + "/Level_0/" + kNames[k], 1, {&f}, kNames[k], d);
            upload(f, *q->f[k], st);
            GUST_HIP(hipStreamSynchronize(st));
        }
        if (z_nd) {
            HostField z;
            z.box = grow(loc_box(g, 7), 1, 1, 1);
            z.v.assign(size_t(z.box.npts()), 0.0);
            read_vismf(d + "/Level_0/z_phys_nd", 1, {&z}, "z_phys_nd", d);
            upload(z, *z_nd, st);
            GUST_HIP(hipStreamSynchronize(st));
        }
        *step = s;
    });
}

}   // extern "C"
