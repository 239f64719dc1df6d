// CIF ingest and supercell construction for the host classes.
//
// CIFReader (cif/cifreader.cpp) reads the cell geometry, the symmetry operations
// (_symmetry_equiv_pos_as_xyz, or _space_group_symop_operation_xyz) and the atom sites
// (label, type symbol, fractional position, occupancy); makeSuperCell (cif/supercell.cpp:16-174)
// rotates the cell so the zone axis [uvw] lies along z (and the optional [abc] along x),
// applies small tilts, tiles it over the requested widths and keeps the atoms inside them;
// processOccupancyList (structure/crystalstructure.cpp:214-235) keeps one atom of each group
// of atoms sharing a site, chosen by one uniform draw against the cumulative occupancies.
// Here the draw comes from Philox4x32-10 keyed by a seed (counter (group, 0x0CC, 0, 3)),
// not from a clock-seeded generator, and site positions are compared with a 1e-6 tolerance
// in fractional coordinates when symmetry copies are merged.
#include <array>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

#include "cltem_host.hpp"

namespace cltem_amd {

namespace {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

Vec3 mul(const Mat3& m, const Vec3& v) {
    return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2], m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}
Mat3 mul(const Mat3& a, const Mat3& b) {
    Mat3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) r[i][j] += a[i][k] * b[k][j];
    return r;
}
Mat3 identity() { return {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
Vec3 normalized(Vec3 a) {
    const double n = std::sqrt(dot(a, a));
    return {a[0] / n, a[1] / n, a[2] / n};
}
bool approx(const Vec3& a, const Vec3& b) {
    // Eigen's isApprox: |a - b| <= 1e-12 * min(|a|, |b|)
    const Vec3 d = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return std::sqrt(dot(d, d)) <= 1e-12 * std::min(std::sqrt(dot(a, a)), std::sqrt(dot(b, b)));
}
Mat3 skew(const Vec3& v) { return {{{0, -v[2], v[1]}, {v[2], 0, -v[0]}, {-v[1], v[0], 0}}}; }

// generateNormalisedRotationMatrix (cif/cifutilities.h:47-76): the rotation taking A onto B
Mat3 align(Vec3 A, Vec3 B) {
    A = normalized(A);
    B = normalized(B);
    if (approx(A, B)) return identity();
    Vec3 v = cross(A, B);
    if (approx(A, {-B[0], -B[1], -B[2]})) {
        Vec3 t = {1.0, 0.0, 0.0};
        if (approx(t, A)) t[2] = 1.0;
        v = cross(A, t);
    }
    const double c = dot(A, B), s2 = dot(v, v);
    const Mat3 k = skew(v), kk = mul(k, k);
    Mat3 r = identity();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] += k[i][j] + kk[i][j] * (1 - c) / s2;
    return r;
}

// generateRotationMatrix (cif/cifutilities.h:79-91): rotation by theta about ax
Mat3 rotation(Vec3 ax, double theta) {
    ax = normalized(ax);
    const Mat3 c = skew(ax), cc = mul(c, c);
    Mat3 r = identity();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] += c[i][j] * std::sin(theta) + cc[i][j] * (1 - std::cos(theta));
    return r;
}

std::string strip_uncertainty(const std::string& s) { return s.substr(0, s.find('(')); }

double number(const std::string& s) {
    const std::string t = strip_uncertainty(s);
    size_t used = 0;
    const double v = std::stod(t, &used);
    if (used != t.size()) throw std::runtime_error("CIF: not a number: " + s);
    return v;
}

// CIF tokens: whitespace separated, '...' / "..." quoted, # comments, ;-delimited text fields
std::vector<std::string> tokenize(std::istream& in) {
    std::vector<std::string> out;
    std::string line;
    bool text = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty() && line[0] == ';') {  // text field delimiter
            text = !text;
            if (!text) out.push_back("<text>");
            continue;
        }
        if (text) continue;
        size_t i = 0;
        while (i < line.size()) {
            const char ch = line[i];
            if (std::isspace((unsigned char)ch)) {
                ++i;
            } else if (ch == '#') {
                break;
            } else if (ch == '\'' || ch == '"') {
                size_t j = i + 1;
                while (j < line.size() && !(line[j] == ch && (j + 1 == line.size() || std::isspace((unsigned char)line[j + 1]))))
                    ++j;
                out.push_back(line.substr(i + 1, j - i - 1));
                i = j + 1;
            } else {
                size_t j = i;
                while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
                out.push_back(line.substr(i, j - i));
                i = j;
            }
        }
    }
    return out;
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// "x", "-y+1/2", "x-y", "+z+0.25": coefficients of x, y, z and the translation
std::array<double, 4> parse_component(const std::string& expr) {
    std::array<double, 4> r{0, 0, 0, 0};
    std::string e;
    for (char c : expr)
        if (!std::isspace((unsigned char)c)) e += (char)std::tolower((unsigned char)c);
    size_t i = 0;
    while (i < e.size()) {
        double sign = 1.0;
        if (e[i] == '+' || e[i] == '-') {
            sign = e[i] == '-' ? -1.0 : 1.0;
            ++i;
        }
        if (i >= e.size()) throw std::runtime_error("CIF: bad symmetry term " + expr);
        if (e[i] == 'x' || e[i] == 'y' || e[i] == 'z') {
            r[e[i] - 'x'] += sign;
            ++i;
            continue;
        }
        size_t j = i;
        while (j < e.size() && (std::isdigit((unsigned char)e[j]) || e[j] == '.')) ++j;
        if (j == i) throw std::runtime_error("CIF: bad symmetry term " + expr);
        double v = std::stod(e.substr(i, j - i));
        if (j < e.size() && e[j] == '/') {
            size_t k = j + 1;
            while (k < e.size() && std::isdigit((unsigned char)e[k])) ++k;
            v /= std::stod(e.substr(j + 1, k - j - 1));
            j = k;
        }
        if (j < e.size() && (e[j] == 'x' || e[j] == 'y' || e[j] == 'z')) {  // "2x"
            r[e[j] - 'x'] += sign * v;
            ++j;
        } else {
            r[3] += sign * v;
        }
        i = j;
    }
    return r;
}

double wrap(double p) {
    p -= std::floor(p);
    if (p >= 1.0 - 1e-12) p = 0.0;  // -0.0 / 1.0 round-off back to the origin
    return p;
}

std::string element_of(const std::string& s) {
    // leading one or two letters, as a symbol: "O2-" -> O, "Sr1" -> Sr, "TI" -> Ti
    std::string sym;
    for (char c : s) {
        if (!std::isalpha((unsigned char)c) || sym.size() == 2) break;
        sym += sym.empty() ? (char)std::toupper((unsigned char)c) : (char)std::tolower((unsigned char)c);
    }
    if (sym.size() == 2) {
        try {
            element_number(sym);
        } catch (const std::exception&) {
            sym.resize(1);
        }
    }
    return sym;
}

}  // namespace

CifStructure read_cif(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error opening .cif file: " + path);
    const std::vector<std::string> tok = tokenize(in);
    CifStructure cif;
    bool have[6] = {false, false, false, false, false, false};
    const char* cell_tags[6] = {"_cell_length_a", "_cell_length_b", "_cell_length_c",
                                "_cell_angle_alpha", "_cell_angle_beta", "_cell_angle_gamma"};
    double* cell_vals[6] = {&cif.a, &cif.b, &cif.c, &cif.alpha, &cif.beta, &cif.gamma};
    std::vector<std::string> sym_ops;
    struct RawSite { std::string label, symbol; double x, y, z, occ; };
    std::vector<RawSite> raw;
    for (size_t i = 0; i < tok.size();) {
        const std::string t = lower(tok[i]);
        if (t == "loop_") {
            size_t j = i + 1;
            std::vector<std::string> heads;
            while (j < tok.size() && tok[j][0] == '_') heads.push_back(lower(tok[j++]));
            size_t k = j;
            while (k < tok.size() && tok[k][0] != '_' && lower(tok[k]) != "loop_" && lower(tok[k]).rfind("data_", 0) != 0)
                ++k;
            const size_t nv = k - j, nh = heads.size();
            if (nh == 0 || nv % nh != 0) throw std::runtime_error("CIF: loop with " + std::to_string(nv) +
                                                                  " values for " + std::to_string(nh) + " tags");
            auto col = [&](const std::string& h) {
                for (size_t c = 0; c < nh; ++c)
                    if (heads[c] == h) return (int)c;
                return -1;
            };
            const int cs = std::max(col("_symmetry_equiv_pos_as_xyz"), col("_space_group_symop_operation_xyz"));
            const int cl = col("_atom_site_label"), ct = col("_atom_site_type_symbol");
            const int cx = col("_atom_site_fract_x"), cy = col("_atom_site_fract_y"), cz = col("_atom_site_fract_z");
            const int co = col("_atom_site_occupancy");
            for (size_t r = 0; r < nv / nh; ++r) {
                auto v = [&](int c) { return tok[j + r * nh + c]; };
                if (cs >= 0) sym_ops.push_back(v(cs));
                if (cx >= 0 && cy >= 0 && cz >= 0) {
                    if (cl < 0 && ct < 0) throw std::runtime_error("CIF: atom sites without label or type symbol");
                    RawSite s;
                    s.label = cl >= 0 ? v(cl) : v(ct);
                    s.symbol = element_of(ct >= 0 ? v(ct) : v(cl));
                    s.x = number(v(cx));
                    s.y = number(v(cy));
                    s.z = number(v(cz));
                    s.occ = (co >= 0 && v(co) != "?" && v(co) != ".") ? number(v(co)) : 1.0;
                    raw.push_back(s);
                }
            }
            i = k;
            continue;
        }
        bool used = false;
        for (int c = 0; c < 6; ++c)
            if (t == cell_tags[c] && i + 1 < tok.size()) {
                *cell_vals[c] = number(tok[i + 1]);
                have[c] = true;
                used = true;
            }
        i += used ? 2 : 1;
    }
    for (int c = 0; c < 6; ++c)
        if (!have[c]) throw std::runtime_error(std::string("CIF: missing ") + cell_tags[c]);
    if (raw.empty()) throw std::runtime_error("CIF: no _atom_site_ block");
    if (sym_ops.empty()) sym_ops.push_back("x,y,z");
    for (const std::string& op : sym_ops) {
        std::vector<std::string> parts;
        std::stringstream ss(op);
        std::string p;
        while (std::getline(ss, p, ',')) parts.push_back(p);
        if (parts.size() != 3) throw std::runtime_error("CIF: symmetry operation needs 3 terms: " + op);
        CifStructure::Op o;
        for (int r = 0; r < 3; ++r) o.row[r] = parse_component(parts[r]);
        cif.ops.push_back(o);
    }
    // sites at the same base position share one site (several elements, occupancies)
    for (const RawSite& s : raw) {
        const int Z = element_number(s.symbol);
        bool merged = false;
        for (auto& site : cif.sites) {
            if (std::fabs(site.fx - s.x) < 1e-6 && std::fabs(site.fy - s.y) < 1e-6 && std::fabs(site.fz - s.z) < 1e-6) {
                site.Z.push_back(Z);
                site.occ.push_back(s.occ);
                merged = true;
                break;
            }
        }
        if (!merged) cif.sites.push_back({s.label, {Z}, {s.occ}, s.x, s.y, s.z, {}});
    }
    // symmetry copies of every site inside the cell (AtomSite::applySymmetry, cif/atomsite.cpp:16-28)
    for (auto& site : cif.sites) {
        for (const auto& o : cif.ops) {
            Vec3 f;
            for (int r = 0; r < 3; ++r)
                f[r] = wrap(o.row[r][0] * site.fx + o.row[r][1] * site.fy + o.row[r][2] * site.fz + o.row[r][3]);
            bool dup = false;
            for (const auto& q : site.positions) {
                double d = 0;
                for (int r = 0; r < 3; ++r) {
                    double dd = std::fabs(q[r] - f[r]);
                    dd = std::min(dd, 1.0 - dd);  // periodic
                    d = std::max(d, dd);
                }
                if (d < 1e-6) dup = true;
            }
            if (!dup) site.positions.push_back(f);
        }
    }
    return cif;
}

AtomList cif_supercell(const CifStructure& cif, const double uvw[3], const double abc[3], const double widths[3],
                       const double tilts_deg[3], uint64_t seed, std::vector<double>* occupancy) {
    // CellGeometry::calculateCartesianBasis (cif/cellgeometry.cpp:18-36)
    const double ar = cif.alpha * (M_PI / 180), br = cif.beta * (M_PI / 180), gr = cif.gamma * (M_PI / 180);
    std::vector<Vec3> basis = {
        {cif.a, 0.0, 0.0},
        {cif.b * std::cos(gr), cif.b * std::sin(gr), 0.0},
        {cif.c * std::cos(br), cif.c * (std::cos(ar) - std::cos(br) * std::cos(gr)) / std::sin(gr),
         cif.c * std::sqrt((std::cos(br - gr) - std::cos(ar)) * (std::cos(ar) - std::cos(br + gr))) / std::sin(gr)}};
    const Vec3 uvw_vec = {uvw[0] * basis[0][0] + uvw[1] * basis[1][0] + uvw[2] * basis[2][0],
                          uvw[0] * basis[0][1] + uvw[1] * basis[1][1] + uvw[2] * basis[2][1],
                          uvw[0] * basis[0][2] + uvw[1] * basis[1][2] + uvw[2] * basis[2][2]};
    const Mat3 za = align(uvw_vec, {0.0, 0.0, 1.0});
    Mat3 xy = identity();
    if (abc[0] != 0 || abc[1] != 0 || abc[2] != 0) {
        Vec3 v = {abc[0] * basis[0][0] + abc[1] * basis[1][0] + abc[2] * basis[2][0],
                  abc[0] * basis[0][1] + abc[1] * basis[1][1] + abc[2] * basis[2][1],
                  abc[0] * basis[0][2] + abc[1] * basis[1][2] + abc[2] * basis[2][2]};
        v = mul(za, v);
        v[2] = 0.0;
        if (std::fabs(v[0]) < 1e-15) v[0] = 0.0;
        if (std::fabs(v[1]) < 1e-15) v[1] = 0.0;
        const Vec3 xdir = {1.0, 0.0, 0.0};
        const Vec3 vn = normalized(v);
        if (approx(xdir, {-vn[0], -vn[1], -vn[2]}))
            xy = rotation({0.0, 0.0, 1.0}, M_PI);
        else
            xy = align(xdir, v);  // supercell.cpp: the negative angle, arguments in this order
    }
    const Mat3 rx = rotation({1, 0, 0}, tilts_deg[0] * M_PI / 180), ry = rotation({0, 1, 0}, tilts_deg[1] * M_PI / 180),
               rz = rotation({0, 0, 1}, tilts_deg[2] * M_PI / 180);
    for (auto& b : basis) b = mul(rz, mul(ry, mul(rx, mul(xy, mul(za, b)))));
    // calculateTiling (supercell.cpp:176-240): the cuboid's corners in cell coordinates
    Mat3 B;  // columns = basis vectors
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) B[r][c] = basis[c][r];
    const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                       B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
    if (std::fabs(det) < 1e-12) throw std::runtime_error("CIF: degenerate cell");
    Mat3 inv;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
            inv[r][c] = (B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1]) / det;
        }
    long mins[3] = {LONG_MAX, LONG_MAX, LONG_MAX}, maxs[3] = {LONG_MIN, LONG_MIN, LONG_MIN};
    for (int corner = 0; corner < 8; ++corner) {
        const Vec3 p = {(corner & 1) ? widths[0] : 0.0, (corner & 2) ? widths[1] : 0.0, (corner & 4) ? widths[2] : 0.0};
        const Vec3 f = mul(inv, p);
        for (int d = 0; d < 3; ++d) {
            const long rnd = (long)((f[d] > 0) - (f[d] < 0)) * (long)std::ceil(std::fabs(f[d]));
            mins[d] = std::min(mins[d], rnd);
            maxs[d] = std::max(maxs[d], rnd);
        }
    }
    AtomList out;
    std::vector<double> occ;
    std::vector<int> group;  // occupancy group of each emitted atom
    int n_groups = 0;
    for (const auto& site : cif.sites)
        for (const auto& f : site.positions) {
            const Vec3 p = {f[0] * basis[0][0] + f[1] * basis[1][0] + f[2] * basis[2][0],
                            f[0] * basis[0][1] + f[1] * basis[1][1] + f[2] * basis[2][1],
                            f[0] * basis[0][2] + f[1] * basis[1][2] + f[2] * basis[2][2]};
            for (long k = mins[2]; k < maxs[2]; ++k)
                for (long j = mins[1]; j < maxs[1]; ++j)
                    for (long i = mins[0]; i < maxs[0]; ++i) {
                        Vec3 q;
                        for (int d = 0; d < 3; ++d) q[d] = p[d] + basis[0][d] * i + basis[1][d] * j + basis[2][d] * k;
                        if (!(q[0] >= 0.0 && q[0] <= widths[0] && q[1] >= 0.0 && q[1] <= widths[1] && q[2] >= 0.0 &&
                              q[2] <= widths[2]))
                            continue;
                        for (size_t e = 0; e < site.Z.size(); ++e) {
                            out.Z.push_back(site.Z[e]);
                            out.x.push_back(q[0]);
                            out.y.push_back(q[1]);
                            out.z.push_back(q[2]);
                            occ.push_back(site.occ[e]);
                            group.push_back(n_groups);
                        }
                        ++n_groups;
                    }
        }
    if (out.size() == 0) throw std::runtime_error("Did not find any atoms inside limits");
    if (occupancy) *occupancy = occ;
    // processOccupancyList: one draw per group, atom kept when the draw falls in its share
    AtomList kept;
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    size_t i = 0;
    while (i < out.size()) {
        size_t j = i;
        while (j < out.size() && group[j] == group[i]) ++j;
        const bool all_one = (j - i == 1) && occ[i] == 1.0;
        double r = 0.0;
        if (!all_one) {
            const uint32_t ctr[4] = {(uint32_t)group[i], 0x0CCu, 0u, 3u};
            uint32_t rr[4];
            philox4x32_10(ctr, key, rr);
            r = ((double)rr[0] + 0.5) / 4294967296.0;
        }
        double total = 0.0;
        for (size_t m = i; m < j; ++m) {
            if (all_one || (r >= total && r < total + occ[m])) {
                kept.Z.push_back(out.Z[m]);
                kept.x.push_back(out.x[m]);
                kept.y.push_back(out.y[m]);
                kept.z.push_back(out.z[m]);
            }
            total += occ[m];
        }
        if (total > 1.0000001) throw std::runtime_error("CIF: occupancies of one site sum to more than 1");
        i = j;
    }
    return kept;
}

}  // namespace cltem_amd
