// plotfile.hpp -- AMReX plotfiles of the modal solution, written and read
// without AMReX: the directory layout and text headers of
// WriteMultiLevelPlotfile (the top-level Header in the "HyperCLaw-V1.1"
// form, Level_<l>/Cell_H in the VisMF version-1 form) and the FAB records of
// Level_<l>/Cell_D_00000 (a "FAB <real descriptor><box> <ncomp>" line, then
// the doubles of each component in turn, x fastest), little-endian IEEE
// doubles.  Used by the reference's PlotFile / ReadPlotFile
// (Core/Solver.h:599-652,922-968): one plotfile per conserved variable q,
// "tstep_<n>_q_<q>_plt", components "<variable name>_<mode>".
#ifndef ADG_HOST_PLOTFILE_HPP
#define ADG_HOST_PLOTFILE_HPP

#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace adg {
namespace plt {

struct IBox {
    std::array<int, 3> lo{{0, 0, 0}}, hi{{0, 0, 0}};
};

// One level of a plotfile: its boxes and, per box, ncomp * cells doubles
// (component slowest, x fastest).
struct Level {
    std::array<int, 3> domain_hi{{0, 0, 0}};   // domain is [0, domain_hi]
    std::array<double, 3> dx{{1, 1, 1}};
    std::vector<IBox> boxes;
    std::vector<std::vector<double>> data;
};

struct Plotfile {
    int dim = 2;
    double time = 0.0;
    int step = 0;
    std::vector<std::string> names;
    std::array<double, 3> prob_lo{{0, 0, 0}}, prob_hi{{1, 1, 1}};
    std::vector<Level> levels;
};

inline std::string ivect(const std::array<int, 3>& v, int dim) {
    std::string s = "(";
    for (int d = 0; d < dim; ++d) s += (d ? "," : "") + std::to_string(v[d]);
    return s + ")";
}
inline std::string box_str(const IBox& b, int dim) {
    std::array<int, 3> z{{0, 0, 0}};
    return "(" + ivect(b.lo, dim) + " " + ivect(b.hi, dim) + " " + ivect(z, dim) + ")";
}
inline long box_cells(const IBox& b, int dim) {
    long n = 1;
    for (int d = 0; d < dim; ++d) n *= (long)(b.hi[d] - b.lo[d] + 1);
    return n;
}

inline void mkdir_p(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
    struct stat st;
    if (::stat(path.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) throw std::runtime_error("cannot create " + path);
}

// Little-endian IEEE double, as AMReX's RealDescriptor names it.
constexpr const char* kRealDesc = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

inline void write(const std::string& dir, const Plotfile& pf) {
    const int D = pf.dim, nc = (int)pf.names.size();
    const int finest = (int)pf.levels.size() - 1;
    mkdir_p(dir);
    for (int l = 0; l <= finest; ++l) {
        const Level& L = pf.levels[(size_t)l];
        const std::string ld = dir + "/Level_" + std::to_string(l);
        mkdir_p(ld);
        std::vector<long> off;
        std::vector<std::vector<double>> mn, mx;
        {
            std::ofstream f(ld + "/Cell_D_00000", std::ios::binary);
            if (!f) throw std::runtime_error("cannot write " + ld + "/Cell_D_00000");
            for (size_t b = 0; b < L.boxes.size(); ++b) {
                off.push_back((long)f.tellp());
                const std::string h = "FAB " + std::string(kRealDesc) + box_str(L.boxes[b], D) + " " +
                                      std::to_string(nc) + "\n";
                f.write(h.data(), (std::streamsize)h.size());
                const std::vector<double>& v = L.data[b];
                f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
                const long np = box_cells(L.boxes[b], D);
                std::vector<double> a(nc), z(nc);
                for (int c = 0; c < nc; ++c) {
                    a[c] = z[c] = v[(size_t)c * np];
                    for (long i = 1; i < np; ++i) {
                        a[c] = std::min(a[c], v[(size_t)c * np + i]);
                        z[c] = std::max(z[c], v[(size_t)c * np + i]);
                    }
                }
                mn.push_back(a);
                mx.push_back(z);
            }
        }
        std::ofstream h(ld + "/Cell_H");
        h.precision(17);
        h << 1 << '\n' << 1 << '\n' << nc << '\n' << 0 << '\n';
        h << '(' << L.boxes.size() << " 0\n";
        for (const IBox& b : L.boxes) h << box_str(b, D) << '\n';
        h << ")\n";
        h << L.boxes.size() << '\n';
        for (long o : off) h << "FabOnDisk: Cell_D_00000 " << o << '\n';
        h << '\n' << L.boxes.size() << ',' << nc << '\n';
        for (const auto& a : mn) {
            for (double v : a) h << v << ',';
            h << '\n';
        }
        h << '\n' << L.boxes.size() << ',' << nc << '\n';
        for (const auto& a : mx) {
            for (double v : a) h << v << ',';
            h << '\n';
        }
    }
    std::ofstream h(dir + "/Header");
    h.precision(17);
    h << "HyperCLaw-V1.1\n" << nc << '\n';
    for (const std::string& n : pf.names) h << n << '\n';
    h << D << '\n' << pf.time << '\n' << finest << '\n';
    for (int d = 0; d < D; ++d) h << pf.prob_lo[d] << ' ';
    h << '\n';
    for (int d = 0; d < D; ++d) h << pf.prob_hi[d] << ' ';
    h << '\n';
    for (int l = 0; l < finest; ++l) h << 2 << ' ';
    h << '\n';
    for (int l = 0; l <= finest; ++l) {
        IBox dom;
        dom.hi = pf.levels[(size_t)l].domain_hi;
        h << box_str(dom, D) << ' ';
    }
    h << '\n';
    for (int l = 0; l <= finest; ++l) h << pf.step << ' ';
    h << '\n';
    for (int l = 0; l <= finest; ++l) {
        for (int d = 0; d < D; ++d) h << pf.levels[(size_t)l].dx[d] << ' ';
        h << '\n';
    }
    h << 0 << '\n' << 0 << '\n';
    for (int l = 0; l <= finest; ++l) {
        const Level& L = pf.levels[(size_t)l];
        h << l << ' ' << L.boxes.size() << ' ' << pf.time << '\n' << pf.step << '\n';
        for (const IBox& b : L.boxes)
            for (int d = 0; d < D; ++d)
                h << pf.prob_lo[d] + b.lo[d] * L.dx[d] << ' ' << pf.prob_lo[d] + (b.hi[d] + 1) * L.dx[d] << '\n';
        h << "Level_" << l << "/Cell\n";
    }
}

inline std::vector<int> parse_ints(const std::string& s) {
    std::vector<int> out;
    std::string cur;
    for (char ch : s) {
        if ((ch >= '0' && ch <= '9') || ch == '-') cur += ch;
        else if (!cur.empty()) { out.push_back(std::stoi(cur)); cur.clear(); }
    }
    if (!cur.empty()) out.push_back(std::stoi(cur));
    return out;
}

inline Plotfile read(const std::string& dir) {
    Plotfile pf;
    std::ifstream h(dir + "/Header");
    if (!h) throw std::runtime_error("cannot read " + dir + "/Header");
    std::string line;
    std::getline(h, line);
    if (line != "HyperCLaw-V1.1") throw std::runtime_error(dir + ": not a HyperCLaw-V1.1 plotfile");
    int nc = 0, finest = 0;
    h >> nc;
    pf.names.resize((size_t)nc);
    for (auto& n : pf.names) h >> n;
    h >> pf.dim >> pf.time >> finest;
    const int D = pf.dim;
    std::getline(h, line);   // end of the finest-level line
    std::getline(h, line);
    {
        std::istringstream is(line);
        for (int d = 0; d < D; ++d) is >> pf.prob_lo[d];
    }
    std::getline(h, line);
    {
        std::istringstream is(line);
        for (int d = 0; d < D; ++d) is >> pf.prob_hi[d];
    }
    std::getline(h, line);   // refinement ratios
    std::getline(h, line);   // domains
    pf.levels.resize((size_t)finest + 1);
    {
        const std::vector<int> v = parse_ints(line);
        for (int l = 0; l <= finest; ++l)
            for (int d = 0; d < D; ++d) pf.levels[(size_t)l].domain_hi[d] = v[(size_t)(l * 3 * D + D + d)];
    }
    for (int l = 0; l <= finest; ++l) h >> pf.step;
    for (int l = 0; l <= finest; ++l)
        for (int d = 0; d < D; ++d) h >> pf.levels[(size_t)l].dx[d];
    int coord, bw;
    h >> coord >> bw;
    for (int l = 0; l <= finest; ++l) {
        int lev, nb, step;
        double tm;
        h >> lev >> nb >> tm >> step;
        for (int k = 0; k < nb * D; ++k) { double a, b; h >> a >> b; }
        std::string path;
        h >> path;
        Level& L = pf.levels[(size_t)l];
        // boxes and offsets from Level_<l>/Cell_H
        std::ifstream c(dir + "/" + path + "_H");
        if (!c) throw std::runtime_error("cannot read " + dir + "/" + path + "_H");
        int vers, how, ncc, ng;
        c >> vers >> how >> ncc >> ng;
        std::getline(c, line);
        std::getline(c, line);   // "(nb 0"
        for (int b = 0; b < nb; ++b) {
            std::getline(c, line);
            const std::vector<int> v = parse_ints(line);
            IBox bx;
            for (int d = 0; d < D; ++d) { bx.lo[d] = v[(size_t)d]; bx.hi[d] = v[(size_t)(D + d)]; }
            L.boxes.push_back(bx);
        }
        std::getline(c, line);   // ")"
        int nfab = 0;
        c >> nfab;
        std::vector<long> off((size_t)nfab);
        std::vector<std::string> file((size_t)nfab);
        for (int b = 0; b < nfab; ++b) {
            std::string tag;
            c >> tag >> file[(size_t)b] >> off[(size_t)b];
        }
        for (int b = 0; b < nfab; ++b) {
            std::ifstream f(dir + "/Level_" + std::to_string(l) + "/" + file[(size_t)b], std::ios::binary);
            f.seekg(off[(size_t)b]);
            std::getline(f, line);
            if (line.compare(0, 4, "FAB ") != 0) throw std::runtime_error("bad FAB record in " + dir);
            std::vector<double> v((size_t)nc * box_cells(L.boxes[(size_t)b], D));
            f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
            if (!f) throw std::runtime_error("short FAB record in " + dir);
            L.data.push_back(std::move(v));
        }
    }
    return pf;
}

}  // namespace plt
}  // namespace adg

#endif
