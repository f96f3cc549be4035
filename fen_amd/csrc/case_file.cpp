// case_file.cpp -- the JSON case file that sets up a run (fen_case_read),
// and the context-free text formats FEN writes: the status line
// (src/navier_stokes.f90:734-748) and grid.json (src/grid.f90:233-265).
//
// FEN itself sets a run up in Fortran source (grid%setup, init_solver,
// module variables, set_timestep) and writes JSON only as output.  The case
// file describes the same setup as data; its "Grid" object has the keys of
// the grid.json that grid%print_json writes, so that file is itself the
// start of a case file.  Host code only: nothing here touches the GPU.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fen.h"

namespace {

// ------------------------------------------------------------------ JSON
struct Value {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;   // in file order

    const Value *get(const std::string &k) const
    {
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class Parser {
  public:
    explicit Parser(const std::string &t) : s(t) {}

    Value document()
    {
        Value v = value();
        ws();
        if (p != s.size()) fail("unexpected text after the top-level value");
        return v;
    }

  private:
    static constexpr int kMaxDepth = 64;   // case files nest 3 deep; this bounds the recursion
    const std::string &s;
    size_t p = 0;
    int depth = 0;

    [[noreturn]] void fail(const std::string &what) const
    {
        int line = 1, col = 1;
        for (size_t i = 0; i < p && i < s.size(); i++) {
            if (s[i] == '\n') { line++; col = 1; }
            else col++;
        }
        throw std::runtime_error("line " + std::to_string(line) + ", column " + std::to_string(col) + ": " + what);
    }

    void ws()
    {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) p++;
    }

    bool lit(const char *w)
    {
        const size_t n = std::strlen(w);
        if (s.compare(p, n, w) == 0) { p += n; return true; }
        return false;
    }

    Value value()
    {
        ws();
        if (p >= s.size()) fail("unexpected end of file");
        if (depth >= kMaxDepth) fail("nesting deeper than " + std::to_string(kMaxDepth) + " levels");
        struct Nest {
            int &d;
            explicit Nest(int &x) : d(x) { d++; }
            ~Nest() { d--; }
        } nest(depth);
        Value v;
        const char ch = s[p];
        if (ch == '{') {
            v.kind = Value::OBJ;
            p++;
            ws();
            if (p < s.size() && s[p] == '}') { p++; return v; }
            for (;;) {
                ws();
                if (p >= s.size() || s[p] != '"') fail("expected a key string");
                const size_t key_at = p;
                std::string k = string();
                for (const auto &kv : v.obj)
                    if (kv.first == k) {
                        p = key_at;   // report the key's own position
                        fail("duplicate key \"" + k + "\"");
                    }
                ws();
                if (p >= s.size() || s[p] != ':') fail("expected ':'");
                p++;
                v.obj.emplace_back(k, value());
                ws();
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == '}') { p++; return v; }
                fail("expected ',' or '}'");
            }
        }
        if (ch == '[') {
            v.kind = Value::ARR;
            p++;
            ws();
            if (p < s.size() && s[p] == ']') { p++; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == ']') { p++; return v; }
                fail("expected ',' or ']'");
            }
        }
        if (ch == '"') { v.kind = Value::STR; v.str = string(); return v; }
        if (lit("true")) { v.kind = Value::BOOL; v.b = true; return v; }
        if (lit("false")) { v.kind = Value::BOOL; v.b = false; return v; }
        if (lit("null")) return v;
        if (ch == '-' || std::isdigit((unsigned char)ch)) { v.kind = Value::NUM; v.num = number(); return v; }
        fail(std::string("unexpected character '") + ch + "'");
    }

    std::string string()
    {
        p++;   // opening quote
        std::string out;
        while (p < s.size() && s[p] != '"') {
            char ch = s[p++];
            if ((unsigned char)ch < 0x20) fail("control character in a string");
            if (ch == '\\') {
                if (p >= s.size()) fail("unterminated escape");
                const char e = s[p++];
                switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                default: fail("unsupported escape (paths and names are plain text)");
                }
            } else {
                out += ch;
            }
        }
        if (p >= s.size()) fail("unterminated string");
        p++;
        return out;
    }

    // JSON number grammar; the value is converted by strtod (correctly rounded).
    double number()
    {
        const size_t b = p;
        if (s[p] == '-') p++;
        if (p >= s.size() || !std::isdigit((unsigned char)s[p])) fail("bad number");
        if (s[p] == '0') p++;
        else while (p < s.size() && std::isdigit((unsigned char)s[p])) p++;
        if (p < s.size() && s[p] == '.') {
            p++;
            if (p >= s.size() || !std::isdigit((unsigned char)s[p])) fail("bad number");
            while (p < s.size() && std::isdigit((unsigned char)s[p])) p++;
        }
        if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
            p++;
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) p++;
            if (p >= s.size() || !std::isdigit((unsigned char)s[p])) fail("bad number");
            while (p < s.size() && std::isdigit((unsigned char)s[p])) p++;
        }
        const std::string t = s.substr(b, p - b);
        return std::strtod(t.c_str(), nullptr);
    }
};

// ------------------------------------------------------- case -> fen_case
[[noreturn]] void bad(const std::string &where, const std::string &what)
{
    throw std::runtime_error(where + ": " + what);
}

void known_keys(const Value &o, const std::string &where, std::initializer_list<const char *> keys)
{
    for (const auto &kv : o.obj) {
        bool ok = false;
        for (const char *k : keys) ok = ok || kv.first == k;
        if (!ok) bad(where, "unknown key \"" + kv.first + "\"");
    }
}

const Value &object(const Value &parent, const char *key, const std::string &where, bool required)
{
    static const Value empty = [] {
        Value v;
        v.kind = Value::OBJ;
        return v;
    }();
    const Value *v = parent.get(key);
    if (!v) {
        if (required) bad(where, std::string("missing object \"") + key + "\"");
        return empty;
    }
    if (v->kind != Value::OBJ) bad(where, std::string("\"") + key + "\" must be an object");
    return *v;
}

double num(const Value &o, const char *key, const std::string &where, double dflt, bool required = false)
{
    const Value *v = o.get(key);
    if (!v) {
        if (required) bad(where, std::string("missing \"") + key + "\"");
        return dflt;
    }
    if (v->kind != Value::NUM) bad(where, std::string("\"") + key + "\" must be a number");
    return v->num;
}

int integer(const Value &o, const char *key, const std::string &where, int dflt, bool required = false)
{
    const double d = num(o, key, where, dflt, required);
    if (d != std::floor(d) || std::fabs(d) > 2147483647.0) bad(where, std::string("\"") + key + "\" must be an integer");
    return (int)d;
}

bool boolean(const Value &o, const char *key, const std::string &where, bool dflt)
{
    const Value *v = o.get(key);
    if (!v) return dflt;
    if (v->kind != Value::BOOL) bad(where, std::string("\"") + key + "\" must be true or false");
    return v->b;
}

void text(const Value &o, const char *key, const std::string &where, char *dst, size_t cap, const char *dflt)
{
    const Value *v = o.get(key);
    std::string t = dflt;
    if (v) {
        if (v->kind != Value::STR) bad(where, std::string("\"") + key + "\" must be a string");
        t = v->str;
    }
    if (t.size() + 1 > cap) bad(where, std::string("\"") + key + "\" is longer than " + std::to_string(cap - 1));
    std::memset(dst, 0, cap);
    std::memcpy(dst, t.data(), t.size());
}

std::vector<double> numbers(const Value &o, const char *key, const std::string &where, size_t n, bool required)
{
    const Value *v = o.get(key);
    if (!v) {
        if (required) bad(where, std::string("missing \"") + key + "\"");
        return {};
    }
    if (v->kind != Value::ARR || (n && v->arr.size() != n))
        bad(where, std::string("\"") + key + "\" must be an array of " + std::to_string(n) + " numbers");
    std::vector<double> out;
    for (const Value &e : v->arr) {
        if (e.kind != Value::NUM) bad(where, std::string("\"") + key + "\" must hold numbers");
        out.push_back(e.num);
    }
    return out;
}

void steps_list(const Value &o, const char *key, const std::string &where, int *dst, int *count)
{
    const std::vector<double> v = numbers(o, key, where, 0, false);
    if (v.size() > FEN_CASE_MAXLIST)
        bad(where, std::string("\"") + key + "\" holds more than " + std::to_string(FEN_CASE_MAXLIST) + " steps");
    *count = (int)v.size();
    for (size_t i = 0; i < v.size(); i++) {
        if (v[i] != std::floor(v[i]) || v[i] < 0) bad(where, std::string("\"") + key + "\" must hold step numbers");
        dst[i] = (int)v[i];
    }
}

// grid%boundary_conditions strings (src/grid.f90:41-46)
int bc_code(const std::string &s, const std::string &where)
{
    if (s == "Periodic") return FEN_BC_PERIODIC;
    if (s == "Wall") return FEN_BC_WALL;
    if (s == "Inflow") return FEN_BC_INFLOW;
    if (s == "Outflow") return FEN_BC_OUTFLOW;
    bad(where, "boundary condition \"" + s + "\" is not Periodic, Wall, Inflow or Outflow");
}

void fill_case(const Value &root, fen_case *c)
{
    if (root.kind != Value::OBJ) bad("case", "the top level must be an object");
    known_keys(root, "case", {"Grid", "Physics", "Timestep", "Precision", "Initial", "Restart", "Run", "Multiphase"});
    std::memset(c, 0, sizeof *c);

    // "Grid": the keys of grid.json (src/grid.f90:246-258) plus the boundaries
    const Value &g = object(root, "Grid", "Grid", true);
    known_keys(g, "Grid", {"Nx", "Ny", "Nz", "origin", "Lx", "Ly", "Lz", "boundary_conditions", "name"});
    c->n[0] = integer(g, "Nx", "Grid", 0, true);
    c->n[1] = integer(g, "Ny", "Grid", 0, true);
    c->n[2] = integer(g, "Nz", "Grid", 1);
    for (int d = 0; d < 3; d++)
        if (c->n[d] < 1) bad("Grid", "Nx, Ny and Nz must be >= 1");
    c->dim = c->n[2] == 1 ? 2 : 3;
    const std::vector<double> o = numbers(g, "origin", "Grid", 3, false);
    for (int d = 0; d < 3; d++) c->origin[d] = o.empty() ? 0.0 : o[d];
    c->L[0] = num(g, "Lx", "Grid", 0.0, true);
    c->L[1] = num(g, "Ly", "Grid", 0.0, true);
    // a 2-D grid has one cell of size delta in z (Lz = Lx Nz / Nx in FEN's 2-D programs)
    c->L[2] = num(g, "Lz", "Grid", c->L[0] / c->n[0] * c->n[2], c->dim == 3);
    for (int d = 0; d < 3; d++)
        if (!(c->L[d] > 0.0)) bad("Grid", "Lx, Ly and Lz must be > 0");
    text(g, "name", "Grid", c->name, sizeof c->name, "grid");
    for (int f = 0; f < 6; f++) c->bc[f] = FEN_BC_PERIODIC;   // grid%setup's default
    if (const Value *b = g.get("boundary_conditions")) {
        const size_t nf = c->dim == 3 ? 6 : 4;
        if (b->kind != Value::ARR || b->arr.size() != nf)
            bad("Grid", "\"boundary_conditions\" must list " + std::to_string(nf) +
                            " faces (left, right, bottom, top" + (nf == 6 ? ", front, back)" : ")"));
        for (size_t f = 0; f < nf; f++) {
            if (b->arr[f].kind != Value::STR) bad("Grid", "boundary conditions are strings");
            c->bc[f] = bc_code(b->arr[f].str, "Grid");
        }
    }

    // "Physics": navier_stokes_mod's module variables and their defaults
    // (src/navier_stokes.f90:17-23, 42)
    const Value &ph = object(root, "Physics", "Physics", false);
    known_keys(ph, "Physics", {"density", "viscosity", "g", "CFL", "constant_CFL"});
    c->density = num(ph, "density", "Physics", 1.0);
    c->viscosity = num(ph, "viscosity", "Physics", 1.0);
    const std::vector<double> gv = numbers(ph, "g", "Physics", 3, false);
    for (int d = 0; d < 3; d++) c->g[d] = gv.empty() ? 0.0 : gv[d];
    c->cfl = num(ph, "CFL", "Physics", 1.0);
    c->constant_cfl = boolean(ph, "constant_CFL", "Physics", false) ? 1 : 0;
    if (!(c->density > 0.0)) bad("Physics", "density must be > 0");
    if (!(c->viscosity >= 0.0)) bad("Physics", "viscosity must be >= 0");

    // "Timestep": a fixed dt, or U for set_timestep(comp_grid, dt, U)
    const Value &ts = object(root, "Timestep", "Timestep", true);
    known_keys(ts, "Timestep", {"dt", "U"});
    c->dt = num(ts, "dt", "Timestep", 0.0);
    c->U = num(ts, "U", "Timestep", 0.0);
    if ((c->dt > 0.0) == (c->U > 0.0)) bad("Timestep", "give exactly one of \"dt\" (> 0) and \"U\" (> 0)");
    if (c->constant_cfl && !(c->U > 0.0)) bad("Timestep", "constant_CFL starts from set_timestep: give \"U\"");

    // "Precision"
    if (const Value *pr = root.get("Precision")) {
        if (pr->kind != Value::STR) bad("Precision", "must be \"FP64\" or \"FP32_REFINE\"");
        if (pr->str == "FP64") c->precision = FEN_FP64;
        else if (pr->str == "FP32_REFINE") c->precision = FEN_FP32_REFINE;
        else bad("Precision", "must be \"FP64\" or \"FP32_REFINE\"");
    }

    // "Initial": raw float64 interiors, Fortran order (scalar%read's format)
    const Value &in = object(root, "Initial", "Initial", false);
    known_keys(in, "Initial", {"u", "v", "w", "p", "vof"});
    static const char *names[5] = {"u", "v", "w", "p", "vof"};
    for (int f = 0; f < 5; f++) text(in, names[f], "Initial", c->init[f], FEN_CASE_PATHLEN, "");
    if (c->dim == 2 && c->init[2][0]) bad("Initial", "a 2-D case has no w");

    // "Restart": the state file of a step (save_state's data/state_<step>.raw)
    c->restart = -1;
    if (const Value *r = root.get("Restart")) {
        if (r->kind != Value::OBJ) bad("Restart", "must be an object");
        known_keys(*r, "Restart", {"step", "dir"});
        c->restart = integer(*r, "step", "Restart", 0, true);
        if (c->restart < 0) bad("Restart", "step must be >= 0");
        text(*r, "dir", "Restart", c->restart_dir, FEN_CASE_PATHLEN, "data");
    }

    // "Run"
    const Value &run = object(root, "Run", "Run", true);
    known_keys(run, "Run", {"steps", "status_every", "save_state", "save_fields", "data_dir", "write_final"});
    c->steps = integer(run, "steps", "Run", 0, true);
    if (c->steps < 0) bad("Run", "steps must be >= 0");
    c->status_every = integer(run, "status_every", "Run", 1);
    if (c->status_every < 0) bad("Run", "status_every must be >= 0");
    steps_list(run, "save_state", "Run", c->save_state, &c->n_save_state);
    steps_list(run, "save_fields", "Run", c->save_fields, &c->n_save_fields);
    text(run, "data_dir", "Run", c->data_dir, FEN_CASE_PATHLEN, "data");
    c->write_final = boolean(run, "write_final", "Run", false) ? 1 : 0;

    // "Multiphase": multiphase_mod's rho_0, rho_1, mu_0, mu_1, sigma
    // (src/multiphase.f90:13-20); the initial interface is the vof file or
    // the level set y - level - amplitude cos(2 pi (x - x0) / wavelength)
    if (const Value *m = root.get("Multiphase")) {
        if (m->kind != Value::OBJ) bad("Multiphase", "must be an object");
        known_keys(*m, "Multiphase", {"rho_0", "rho_1", "mu_0", "mu_1", "sigma", "interface"});
        c->multiphase = 1;
        c->rho0 = num(*m, "rho_0", "Multiphase", 0.0, true);
        c->rho1 = num(*m, "rho_1", "Multiphase", 0.0, true);
        c->mu0 = num(*m, "mu_0", "Multiphase", 0.0, true);
        c->mu1 = num(*m, "mu_1", "Multiphase", 0.0, true);
        c->sigma = num(*m, "sigma", "Multiphase", 0.0);
        if (!(c->rho0 > 0.0 && c->rho1 > 0.0)) bad("Multiphase", "densities must be > 0");
        if (c->dim != 2) bad("Multiphase", "the two-phase mode is 2-D");
        const Value &itf = object(*m, "interface", "Multiphase", false);
        known_keys(itf, "Multiphase.interface", {"level", "amplitude", "wavelength"});
        c->iface[0] = num(itf, "level", "Multiphase.interface", 0.0, !c->init[4][0]);
        c->iface[1] = num(itf, "amplitude", "Multiphase.interface", 0.0);
        c->iface[2] = num(itf, "wavelength", "Multiphase.interface", c->L[0]);
        if (!(c->iface[2] > 0.0)) bad("Multiphase.interface", "wavelength must be > 0");
    } else if (c->init[4][0]) {
        bad("Initial", "a vof needs the \"Multiphase\" object");
    }
}

void copy_err(char *err, size_t len, const std::string &msg)
{
    if (!err || len == 0) return;
    std::snprintf(err, len, "%s", msg.c_str());
}

}  // namespace

// Fortran Ew.d edit descriptor (0.ddddE+xx), right-justified in w columns.
std::string fen_fortran_e(double x, int w, int d)
{
    char buf[64];
    if (std::isnan(x)) { std::snprintf(buf, sizeof buf, "%*s", w, "NaN"); return buf; }
    if (std::isinf(x)) { std::snprintf(buf, sizeof buf, "%*s", w, x > 0 ? "Infinity" : "-Infinity"); return buf; }
    std::string body;
    if (x == 0.0) {
        body = std::string(std::signbit(x) ? "-" : "") + "0." + std::string(d, '0') + "E+00";
    } else {
        char tmp[64];
        std::snprintf(tmp, sizeof tmp, "%.*E", d - 1, std::fabs(x));   // d.ddddE+xx
        const std::string s(tmp);
        const size_t epos = s.find('E');
        const int e = std::atoi(s.c_str() + epos + 1) + 1;
        const std::string digits = s.substr(0, 1) + s.substr(2, epos - 2);
        char ex[16];
        if (std::abs(e) <= 99) std::snprintf(ex, sizeof ex, "E%c%02d", e < 0 ? '-' : '+', std::abs(e));
        else std::snprintf(ex, sizeof ex, "%c%03d", e < 0 ? '-' : '+', std::abs(e));
        body = std::string(x < 0 ? "-" : "") + "0." + digits + ex;
    }
    std::snprintf(buf, sizeof buf, "%*s", w, body.c_str());
    return buf;
}

extern "C" {

int fen_case_read(const char *path, fen_case *out, char *err, size_t errlen)
{
    if (!path || !out) return FEN_ERR_ARG;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        copy_err(err, errlen, std::string("cannot open case file ") + path);
        return FEN_ERR_ARG;
    }
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
    std::fclose(fp);
    try {
        Parser ps(text);
        const Value root = ps.document();
        fen_case c;
        fill_case(root, &c);
        *out = c;
    } catch (const std::exception &e) {
        copy_err(err, errlen, std::string(path) + ": " + e.what());
        return FEN_ERR_ARG;
    }
    copy_err(err, errlen, "");
    return FEN_OK;
}

int fen_status_text(int step, double time, double dt, double maxdiv, double maxcfl, char *buf, size_t len)
{
    if (!buf || len == 0) return FEN_ERR_ARG;
    // (A6,I7,1x,A6,E13.6,1x,A4,E13.6,1x,A8,E13.6,1x,A9,1x,E13.6), src/navier_stokes.f90:746
    char ib[32];
    std::snprintf(ib, sizeof ib, "%7d", step);
    std::string s = std::string("step: ") + (std::strlen(ib) > 7 ? "*******" : ib);
    s += " time: " + fen_fortran_e(time, 13, 6);
    s += " dt: " + fen_fortran_e(dt, 13, 6);
    s += " maxdiv: " + fen_fortran_e(maxdiv, 13, 6);
    s += "  maxCFL:  " + fen_fortran_e(maxcfl, 13, 6);
    std::snprintf(buf, len, "%s", s.c_str());
    return s.size() < len ? FEN_OK : FEN_ERR_ARG;
}

int fen_grid_json_text(int nx, int ny, int nz, const double *x0, double lx, double ly, double lz, char *buf,
                       size_t len)
{
    if (!buf || len == 0) return FEN_ERR_ARG;
    const double o[3] = {x0 ? x0[0] : 0.0, x0 ? x0[1] : 0.0, x0 ? x0[2] : 0.0};
    auto i7 = [](int v) {
        char b[32];
        std::snprintf(b, sizeof b, "%7d", v);
        return std::strlen(b) > 7 ? std::string("*******") : std::string(b);
    };
    // src/grid.f90:246-258: (A1) (4x,A9) (8x,A6,1x,I7,A1) x3
    // (8x,A11,1x,E16.8,A1,E16.8,A1,E16.8,A2) (8x,A6,1x,E16.8,A1) x2
    // (8x,A6,1x,E16.8) (4x,A3) (A1)
    std::string s = "{\n";
    s += "    \"Grid\": {\n";
    s += "        \"Nx\":  " + i7(nx) + ",\n";
    s += "        \"Ny\":  " + i7(ny) + ",\n";
    s += "        \"Nz\":  " + i7(nz) + ",\n";
    s += "        \"origin\": [ " + fen_fortran_e(o[0], 16, 8) + "," + fen_fortran_e(o[1], 16, 8) + "," +
         fen_fortran_e(o[2], 16, 8) + "],\n";
    s += "        \"Lx\":  " + fen_fortran_e(lx, 16, 8) + ",\n";
    s += "        \"Ly\":  " + fen_fortran_e(ly, 16, 8) + ",\n";
    s += "        \"Lz\":  " + fen_fortran_e(lz, 16, 8) + "\n";
    s += "      }\n";
    s += "}\n";
    std::snprintf(buf, len, "%s", s.c_str());
    return s.size() < len ? FEN_OK : FEN_ERR_ARG;
}

}  // extern "C"
