// gdp_mps.cpp - MPS reader for the propagation entry points.
//
// Reference counterpart: the python-mip based reader of
// fileReader/readerInterface.py:52-145 (PythonMIPReader: n, m, nnz, variable
// bounds, lhs/rhs by constraint sense, the constraint matrix in CSR by
// constraint and variable order, SCIP variable types B=0, I=1, C=3).
// Semantics of fixed-format / free-format MPS as read here:
//   ROWS     N rows are objectives or free rows and are not constraints; L,
//            G, E rows in file order are the constraints (lhs = -inf for L,
//            rhs = +inf for G, lhs = rhs for E)
//   COLUMNS  variables in order of first appearance; MARKER ... 'INTORG' /
//            'INTEND' bracket integer variables; repeated (row, column)
//            entries are summed, zero entries are dropped
//   RHS      rhs of L/E rows, lhs of G/E rows (entries on N rows ignored)
//   RANGES   L: [rhs - |R|, rhs], G: [rhs, rhs + |R|], E: R > 0 -> [rhs,
//            rhs + R], R < 0 -> [rhs + R, rhs]
//   BOUNDS   UP (lb becomes -inf when the value is negative and lb is 0), LO,
//            FX, FR, MI, PL, BV, LI, UI
// Default bounds are [0, +inf) for every variable, integer ones included.
// Infinite bounds are returned as +-1e20 (GDP_INF).  An integer variable
// with bounds [0, 1] gets type BINARY (0), other integers INTEGER (1),
// continuous variables CONTINUOUS (3).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gdp_ext.h"
#include "gdp_engine.h"

namespace {

constexpr double kMpsInf = 1e20;

std::vector<std::string> fields(const std::string &line) {
    std::vector<std::string> f;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) i++;
        size_t j = i;
        while (j < line.size() && line[j] != ' ' && line[j] != '\t' && line[j] != '\r') j++;
        if (j > i) f.push_back(line.substr(i, j - i));
        i = j;
    }
    return f;
}

double number(const std::string &s) {
    errno = 0;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE) throw std::runtime_error("bad number '" + s + "'");
    return v;
}

double clamp_inf(double v) {
    if (v >= kMpsInf) return kMpsInf;
    if (v <= -kMpsInf) return -kMpsInf;
    return v;
}

struct Model {
    std::string name;
    std::vector<char> sense;                 // 'L', 'G', 'E' per constraint
    std::vector<double> rhs_val, range_val;  // RHS and RANGES values
    std::vector<char> has_range;
    std::unordered_map<std::string, int> row_index;  // -1: objective / free row
    std::unordered_map<std::string, int> col_index;
    std::vector<std::map<int, double>> cols;  // per variable: constraint -> coefficient
    std::vector<char> integer;
    std::vector<double> lb, ub;
};

Model parse(const char *text, size_t len) {
    Model M;
    std::string section;
    bool in_int = false;
    size_t pos = 0;
    int lineno = 0;
    while (pos < len) {
        size_t e = pos;
        while (e < len && text[e] != '\n') e++;
        std::string line(text + pos, e - pos);
        pos = e + 1;
        lineno++;
        if (line.empty() || line[0] == '*') continue;
        const bool header = line[0] != ' ' && line[0] != '\t';
        std::vector<std::string> f = fields(line);
        if (f.empty()) continue;
        try {
            if (header) {
                section = f[0];
                if (section == "NAME") M.name = f.size() > 1 ? f[1] : "";
                else if (section == "ENDATA") break;
                else if (section != "ROWS" && section != "COLUMNS" && section != "RHS" && section != "RANGES" &&
                         section != "BOUNDS" && section != "OBJSENSE" && section != "OBJSENSE MAX")
                    throw std::runtime_error("unsupported section " + section);
                continue;
            }
            if (section == "ROWS") {
                if (f.size() < 2) throw std::runtime_error("short ROWS line");
                const char t = f[0][0];
                if (t == 'N') {
                    M.row_index[f[1]] = -1;
                } else if (t == 'L' || t == 'G' || t == 'E') {
                    M.row_index[f[1]] = static_cast<int>(M.sense.size());
                    M.sense.push_back(t);
                    M.rhs_val.push_back(0.0);
                    M.range_val.push_back(0.0);
                    M.has_range.push_back(0);
                } else {
                    throw std::runtime_error("unknown row type " + f[0]);
                }
            } else if (section == "COLUMNS") {
                if (f.size() >= 3 && f[1] == "'MARKER'") {
                    if (f[2] == "'INTORG'") in_int = true;
                    else if (f[2] == "'INTEND'") in_int = false;
                    continue;
                }
                if (f.size() < 3 || f.size() % 2 == 0) throw std::runtime_error("bad COLUMNS line");
                auto it = M.col_index.find(f[0]);
                int j;
                if (it == M.col_index.end()) {
                    j = static_cast<int>(M.cols.size());
                    M.col_index[f[0]] = j;
                    M.cols.emplace_back();
                    M.integer.push_back(in_int ? 1 : 0);
                    M.lb.push_back(0.0);
                    M.ub.push_back(kMpsInf);
                } else {
                    j = it->second;
                }
                for (size_t k = 1; k + 1 < f.size(); k += 2) {
                    auto r = M.row_index.find(f[k]);
                    if (r == M.row_index.end()) throw std::runtime_error("unknown row " + f[k]);
                    if (r->second < 0) continue;  // objective or free row
                    M.cols[j][r->second] += number(f[k + 1]);
                }
            } else if (section == "RHS" || section == "RANGES") {
                // optional set name: an odd number of fields
                size_t k = f.size() % 2 == 1 ? 1 : 0;
                for (; k + 1 < f.size(); k += 2) {
                    auto r = M.row_index.find(f[k]);
                    if (r == M.row_index.end()) throw std::runtime_error("unknown row " + f[k]);
                    if (r->second < 0) continue;
                    const double v = number(f[k + 1]);
                    if (section == "RHS") {
                        M.rhs_val[r->second] = v;
                    } else {
                        M.range_val[r->second] = v;
                        M.has_range[r->second] = 1;
                    }
                }
            } else if (section == "BOUNDS") {
                if (f.size() < 3) throw std::runtime_error("short BOUNDS line");
                const std::string &t = f[0];
                const bool no_value = t == "FR" || t == "MI" || t == "PL" || t == "BV";
                // with a bound-set name the column is the third field
                const std::string &cname = (no_value ? f.size() >= 3 : f.size() >= 4) ? f[2] : f[1];
                auto c = M.col_index.find(cname);
                if (c == M.col_index.end()) throw std::runtime_error("unknown column " + cname);
                const int j = c->second;
                const double v = no_value ? 0.0 : number(f.back());
                if (t == "UP") {
                    if (v < 0 && M.lb[j] == 0.0) M.lb[j] = -kMpsInf;
                    M.ub[j] = clamp_inf(v);
                } else if (t == "LO") {
                    M.lb[j] = clamp_inf(v);
                } else if (t == "FX") {
                    M.lb[j] = clamp_inf(v);
                    M.ub[j] = clamp_inf(v);
                } else if (t == "FR") {
                    M.lb[j] = -kMpsInf;
                    M.ub[j] = kMpsInf;
                } else if (t == "MI") {
                    M.lb[j] = -kMpsInf;
                } else if (t == "PL") {
                    M.ub[j] = kMpsInf;
                } else if (t == "BV") {
                    M.integer[j] = 1;
                    M.lb[j] = 0.0;
                    M.ub[j] = 1.0;
                } else if (t == "LI") {
                    M.integer[j] = 1;
                    M.lb[j] = clamp_inf(v);
                } else if (t == "UI") {
                    M.integer[j] = 1;
                    if (v < 0 && M.lb[j] == 0.0) M.lb[j] = -kMpsInf;
                    M.ub[j] = clamp_inf(v);
                } else {
                    throw std::runtime_error("unsupported bound type " + t);
                }
            }
        } catch (const std::exception &ex) {
            throw std::runtime_error("MPS line " + std::to_string(lineno) + ": " + ex.what());
        }
    }
    return M;
}

template <typename T>
T *copy_out(const std::vector<T> &v) {
    T *p = static_cast<T *>(std::malloc(sizeof(T) * std::max<size_t>(1, v.size())));
    if (!p) throw std::runtime_error("out of memory");
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" {

GDP_RETCODE gdpParseMPS(const char *text, const size_t len, gdp_mps *out) {
    if (!text || !out) return GDP_ERROR;
    std::memset(out, 0, sizeof(*out));
    try {
        Model M = parse(text, len);
        const int m = static_cast<int>(M.sense.size());
        const int n = static_cast<int>(M.cols.size());
        // constraint matrix by rows (variables in column order within a row)
        std::vector<int> row_cnt(m + 1, 0);
        for (int j = 0; j < n; j++)
            for (const auto &kv : M.cols[j])
                if (kv.second != 0.0) row_cnt[kv.first + 1]++;
        std::vector<int> row_ptr(m + 1, 0);
        for (int i = 0; i < m; i++) row_ptr[i + 1] = row_ptr[i] + row_cnt[i + 1];
        const int nnz = row_ptr[m];
        std::vector<int> col(nnz);
        std::vector<double> val(nnz);
        std::vector<int> fill(row_ptr.begin(), row_ptr.end() - 1);
        for (int j = 0; j < n; j++)
            for (const auto &kv : M.cols[j])
                if (kv.second != 0.0) {
                    col[fill[kv.first]] = j;
                    val[fill[kv.first]] = kv.second;
                    fill[kv.first]++;
                }
        std::vector<double> lhs(m), rhs(m);
        for (int i = 0; i < m; i++) {
            const double b = M.rhs_val[i], R = M.range_val[i];
            switch (M.sense[i]) {
                case 'L':
                    lhs[i] = M.has_range[i] ? b - std::fabs(R) : -kMpsInf;
                    rhs[i] = b;
                    break;
                case 'G':
                    lhs[i] = b;
                    rhs[i] = M.has_range[i] ? b + std::fabs(R) : kMpsInf;
                    break;
                default:
                    lhs[i] = M.has_range[i] && R < 0 ? b + R : b;
                    rhs[i] = M.has_range[i] && R > 0 ? b + R : b;
                    break;
            }
        }
        std::vector<int> vt(n);
        for (int j = 0; j < n; j++)
            vt[j] = !M.integer[j] ? 3 : (M.lb[j] == 0.0 && M.ub[j] == 1.0 ? 0 : 1);
        out->n_cons = m;
        out->n_vars = n;
        out->nnz = nnz;
        out->row_ptrs = copy_out(row_ptr);
        out->col_indices = copy_out(col);
        out->vals = copy_out(val);
        out->lhss = copy_out(lhs);
        out->rhss = copy_out(rhs);
        out->lbs = copy_out(M.lb);
        out->ubs = copy_out(M.ub);
        out->vartypes = copy_out(vt);
        std::strncpy(out->name, M.name.c_str(), sizeof(out->name) - 1);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        gdpMpsFree(out);
        gdp::set_last_error(std::string("gdpParseMPS: ") + e.what());
        std::fprintf(stderr, "gdpParseMPS: %s\n", e.what());
        return GDP_ERROR;
    }
}

void gdpMpsFree(gdp_mps *p) {
    if (!p) return;
    std::free(p->row_ptrs);
    std::free(p->col_indices);
    std::free(p->vals);
    std::free(p->lhss);
    std::free(p->rhss);
    std::free(p->lbs);
    std::free(p->ubs);
    std::free(p->vartypes);
    p->row_ptrs = p->col_indices = p->vartypes = nullptr;
    p->vals = p->lhss = p->rhss = p->lbs = p->ubs = nullptr;
}

}  // extern "C"
