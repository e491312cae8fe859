// MPS reader: the whole file is read once; records are tokenized in place; COLUMNS entries are
// kept as (row, column, value) triplets and bucketed by row with a stable counting sort, so each
// constraint lists its entries in file order, as the per-row accumulation of the reference
// does (src/parsers/mps_reader.cpp:65-355).  Semantics in include/parsers/mps_reader.hpp.
#include "parsers/mps_reader.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

namespace sankhya {
namespace parsers {

namespace {

[[noreturn]] void fail(const std::string& what) {
    throw std::runtime_error("MPS parser error: " + what);
}

std::string at_line(long line) { return " (line " + std::to_string(line) + ")"; }

// whitespace-separated fields of one record
void fields(std::string_view s, std::vector<std::string_view>& out) {
    out.clear();
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
        const size_t b = i;
        while (i < s.size() && !std::isspace(static_cast<unsigned char>(s[i]))) ++i;
        if (i > b) out.push_back(s.substr(b, i - b));
    }
}

// a number whose leading characters parse (the rest of the field is ignored, as a stream read
// of a double would); an empty parse throws
Float number(std::string_view f, const char* where, long line) {
    const std::string t(f);
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str()) fail(std::string("invalid numeric value in ") + where + at_line(line));
    return v;
}

struct Row {
    char kind;            // N L G E
    Float rhs = 0.0;
    bool rhs_set = false;
};

struct Col {
    Float cost = 0.0, lo = 0.0, hi = math::kInfinity;
    VariableType type = VariableType::Continuous;
};

enum class Sec { None, Name, Rows, Columns, Rhs, Bounds };

}  // namespace

core::Model read_mps(const std::string& filename) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) fail("could not open file " + filename);
    std::ostringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();

    std::vector<Row> rows;
    std::unordered_map<std::string, int> row_id;
    int objective = -1;
    std::vector<Col> cols;
    std::unordered_map<std::string, int> col_id;
    std::vector<int> tr, tc;          // COLUMNS triplets
    std::vector<Float> tv;
    bool integer_block = false;
    std::string rhs_vector, bound_vector;
    Sec sec = Sec::None;
    std::vector<std::string_view> f;

    long line = 0;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string_view rec(text.data() + pos, eol - pos);
        pos = eol + 1;
        ++line;
        if (!rec.empty() && rec.back() == '\r') rec.remove_suffix(1);
        if (rec.empty() || rec[0] == '*') continue;

        if (!std::isspace(static_cast<unsigned char>(rec[0]))) {   // section header
            fields(rec, f);
            if (f.empty()) continue;
            const std::string_view h = f[0];
            if (h == "NAME") sec = Sec::Name;
            else if (h == "ROWS") sec = Sec::Rows;
            else if (h == "COLUMNS") sec = Sec::Columns;
            else if (h == "RHS") sec = Sec::Rhs;
            else if (h == "BOUNDS") sec = Sec::Bounds;
            else if (h == "ENDATA") break;
            else fail("unsupported section '" + std::string(h) + "'" + at_line(line));
            continue;
        }
        fields(rec, f);
        if (f.empty()) continue;

        switch (sec) {
        case Sec::Rows: {
            if (f.size() < 2) fail("malformed ROWS record" + at_line(line));
            const char k = f[0][0];
            if (k != 'N' && k != 'L' && k != 'G' && k != 'E')
                fail("unknown row type '" + std::string(1, k) + "'" + at_line(line));
            if (k == 'N' && objective < 0) objective = static_cast<int>(rows.size());
            row_id[std::string(f[1])] = static_cast<int>(rows.size());
            rows.push_back(Row{k});
            break;
        }
        case Sec::Columns: {
            if (f.size() < 3) fail("malformed COLUMNS record" + at_line(line));
            if (f[1] == "'MARKER'") {
                if (f[2] == "'INTORG'") integer_block = true;
                else if (f[2] == "'INTEND'") integer_block = false;
                break;
            }
            const std::string name(f[0]);
            auto it = col_id.find(name);
            if (it == col_id.end()) {
                it = col_id.emplace(name, static_cast<int>(cols.size())).first;
                Col c;
                c.type = integer_block ? VariableType::Integer : VariableType::Continuous;
                cols.push_back(c);
            }
            const int j = it->second;
            for (size_t q = 1; q < f.size(); q += 2) {
                if (q + 1 >= f.size()) fail("missing value in COLUMNS" + at_line(line));
                const Float v = number(f[q + 1], "COLUMNS", line);
                const auto r = row_id.find(std::string(f[q]));
                if (r == row_id.end()) fail("unknown row '" + std::string(f[q]) + "' in COLUMNS" + at_line(line));
                if (r->second == objective) {
                    cols[static_cast<size_t>(j)].cost = v;
                } else {
                    tr.push_back(r->second);
                    tc.push_back(j);
                    tv.push_back(v);
                }
            }
            break;
        }
        case Sec::Rhs: {
            if (f.size() < 3) fail("malformed RHS record" + at_line(line));
            if (rhs_vector.empty()) rhs_vector = std::string(f[0]);
            else if (f[0] != rhs_vector) break;                       // later vectors ignored
            for (size_t q = 1; q < f.size(); q += 2) {
                if (q + 1 >= f.size()) fail("missing value in RHS" + at_line(line));
                const Float v = number(f[q + 1], "RHS", line);
                const auto r = row_id.find(std::string(f[q]));
                if (r == row_id.end()) fail("unknown row '" + std::string(f[q]) + "' in RHS" + at_line(line));
                rows[static_cast<size_t>(r->second)].rhs = v;
                rows[static_cast<size_t>(r->second)].rhs_set = true;
            }
            break;
        }
        case Sec::Bounds: {
            if (f.size() < 3) fail("malformed BOUNDS record" + at_line(line));
            if (bound_vector.empty()) bound_vector = std::string(f[1]);
            else if (f[1] != bound_vector) break;                     // later vectors ignored
            const auto c = col_id.find(std::string(f[2]));
            if (c == col_id.end()) fail("unknown column '" + std::string(f[2]) + "' in BOUNDS" + at_line(line));
            Col& col = cols[static_cast<size_t>(c->second)];
            const std::string_view t = f[0];
            auto value = [&]() {
                if (f.size() < 4) fail(std::string(t) + " bound without a value" + at_line(line));
                return number(f[3], "BOUNDS", line);
            };
            if (t == "LO") col.lo = value();
            else if (t == "UP") col.hi = value();
            else if (t == "FX") col.lo = col.hi = value();
            else if (t == "FR") { col.lo = -math::kInfinity; col.hi = math::kInfinity; }
            else if (t == "MI") col.lo = -math::kInfinity;
            else if (t == "PL") col.hi = math::kInfinity;
            else if (t == "BV") { col.type = VariableType::Binary; col.lo = 0.0; col.hi = 1.0; }
            else if (t == "LI") { col.type = VariableType::Integer; col.lo = value(); }
            else if (t == "UI") { col.type = VariableType::Integer; col.hi = value(); }
            else fail("unsupported bound type '" + std::string(t) + "'" + at_line(line));
            break;
        }
        default:
            break;                                                     // NAME record, or data before a section
        }
    }

    // constraints: rows other than N, in file order; their entries bucketed stably by row
    const int nrows = static_cast<int>(rows.size());
    std::vector<int> start(static_cast<size_t>(nrows) + 1, 0);
    for (int r : tr) ++start[static_cast<size_t>(r) + 1];
    for (int r = 0; r < nrows; ++r) start[static_cast<size_t>(r) + 1] += start[static_cast<size_t>(r)];
    std::vector<int> fill(start.begin(), start.end() - 1), bc(tr.size());
    std::vector<Float> bv(tr.size());
    for (size_t e = 0; e < tr.size(); ++e) {
        const int p = fill[static_cast<size_t>(tr[e])]++;
        bc[static_cast<size_t>(p)] = tc[e];
        bv[static_cast<size_t>(p)] = tv[e];
    }

    core::Model model;
    for (const Col& c : cols) model.add_variable(c.cost, c.lo, c.hi, c.type);
    std::vector<Index> slack(static_cast<size_t>(nrows), -1);
    Index next = static_cast<Index>(cols.size());
    for (int r = 0; r < nrows; ++r) {
        const char k = rows[static_cast<size_t>(r)].kind;
        if (k == 'L' || k == 'G') {
            slack[static_cast<size_t>(r)] = next++;
            model.add_variable(0.0, 0.0, math::kInfinity, VariableType::Continuous);
        }
    }
    std::vector<Index> rc;
    std::vector<Float> rv;
    for (int r = 0; r < nrows; ++r) {
        const Row& row = rows[static_cast<size_t>(r)];
        if (row.kind == 'N') {
            if (row.rhs_set && row.rhs != 0.0)
                fail("a constant objective offset (RHS on an N row) is not supported");
            continue;
        }
        rc.assign(bc.begin() + start[static_cast<size_t>(r)], bc.begin() + start[static_cast<size_t>(r) + 1]);
        rv.assign(bv.begin() + start[static_cast<size_t>(r)], bv.begin() + start[static_cast<size_t>(r) + 1]);
        if (row.kind != 'E') {
            rc.push_back(slack[static_cast<size_t>(r)]);
            rv.push_back(row.kind == 'L' ? 1.0 : -1.0);
        }
        model.add_constraint(rc, rv, row.rhs);
    }
    model.sense = OptimizationSense::Minimize;
    model.finalize();
    return model;
}

}  // namespace parsers
}  // namespace sankhya
