// General CharmmGBMVForce expressions on the device: expression parser, HIP code generation
// over forward-mode dual numbers, and hiprtc compilation.
//
// Reference counterparts: the Reference platform evaluates the expressions with Lepton and
// symbolic derivatives (CharmmReferenceGBMV.cpp:150-204 calculateIxn, :206-227 values,
// :229-266 single-particle terms, :268-348 pair terms, :350-446 chain rule); the CUDA
// platform turns them into generated CUDA source (CharmmCudaKernels.cpp:228-827).  Here one
// evaluation of an expression over dual numbers gives its value and every derivative the
// chain rule needs, so no symbolic derivative is formed.
#include "gbi_custom.h"

#include <hip/hiprtc.h>

#include <cctype>
#include <cmath>
#include <iomanip>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace gbi {

namespace {

// ---------------------------------------------------------------------------
// Expression parser: + - * / ^ (right associative, above unary minus), unary minus,
// parentheses, numbers, names and the functions below.  Emits HIP source over the dual type D.
// ---------------------------------------------------------------------------

struct Token {
    enum Kind { NUM, NAME, OP, END } kind;
    std::string text;
};

std::vector<Token> tokenize(const std::string& s) {
    std::vector<Token> out;
    size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (std::isspace((unsigned char)c)) {
            ++i;
        } else if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.')) ++j;
            if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
                size_t k = j + 1;
                if (k < s.size() && (s[k] == '+' || s[k] == '-')) ++k;
                if (k < s.size() && std::isdigit((unsigned char)s[k])) {
                    j = k;
                    while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
                }
            }
            out.push_back({Token::NUM, s.substr(i, j - i)});
            i = j;
        } else if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
            out.push_back({Token::NAME, s.substr(i, j - i)});
            i = j;
        } else if (std::string("+-*/^(),").find(c) != std::string::npos) {
            out.push_back({Token::OP, std::string(1, c)});
            ++i;
        } else {
            throw std::runtime_error(std::string("unexpected character '") + c + "' in expression \"" + s + "\"");
        }
    }
    out.push_back({Token::END, ""});
    return out;
}

// Functions: name -> (arity, device function over D)
const std::map<std::string, std::pair<int, const char*>>& functions() {
    static const std::map<std::string, std::pair<int, const char*>> f = {
        {"sqrt", {1, "d_sqrt"}},   {"exp", {1, "d_exp"}},     {"log", {1, "d_log"}},   {"sin", {1, "d_sin"}},
        {"cos", {1, "d_cos"}},     {"tan", {1, "d_tan"}},     {"asin", {1, "d_asin"}}, {"acos", {1, "d_acos"}},
        {"atan", {1, "d_atan"}},   {"sinh", {1, "d_sinh"}},   {"cosh", {1, "d_cosh"}}, {"tanh", {1, "d_tanh"}},
        {"erf", {1, "d_erf"}},     {"erfc", {1, "d_erfc"}},   {"abs", {1, "d_abs"}},   {"step", {1, "d_step"}},
        {"delta", {1, "d_delta"}}, {"square", {1, "d_square"}}, {"cube", {1, "d_cube"}}, {"recip", {1, "d_recip"}},
        {"min", {2, "d_min"}},     {"max", {2, "d_max"}},     {"select", {3, "d_select"}}};
    return f;
}

class Parser {
public:
    // tabs: tabulated function name -> (device function, number of arguments)
    Parser(const std::string& expr, const std::map<std::string, std::string>& vars,
           const std::map<std::string, std::pair<std::string, int>>& tabs)
        : src_(expr), toks_(tokenize(expr)), vars_(vars), tabs_(tabs) {}

    std::string parse() {
        std::string e = sum();
        if (toks_[pos_].kind != Token::END) fail("unexpected '" + toks_[pos_].text + "'");
        return e;
    }

private:
    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error(what + " in expression \"" + src_ + "\"");
    }
    bool is_op(const char* op) const { return toks_[pos_].kind == Token::OP && toks_[pos_].text == op; }
    void expect(const char* op) {
        if (!is_op(op)) fail(std::string("expected '") + op + "'");
        ++pos_;
    }
    std::string sum() {
        std::string a = product();
        while (is_op("+") || is_op("-")) {
            const std::string op = toks_[pos_++].text;
            a = "(" + a + " " + op + " " + product() + ")";
        }
        return a;
    }
    std::string product() {
        std::string a = unary();
        while (is_op("*") || is_op("/")) {
            const std::string op = toks_[pos_++].text;
            a = "(" + a + " " + op + " " + unary() + ")";
        }
        return a;
    }
    std::string unary() {
        if (is_op("-")) {
            ++pos_;
            return "(-" + unary() + ")";
        }
        if (is_op("+")) {
            ++pos_;
            return unary();
        }
        return power();
    }
    std::string power() {
        std::string a = primary();
        if (is_op("^")) {
            ++pos_;
            // a numeric exponent (optionally negative) differentiates without log(a)
            if (toks_[pos_].kind == Token::NUM ||
                (is_op("-") && toks_[pos_ + 1].kind == Token::NUM &&
                 !(toks_[pos_ + 2].kind == Token::OP && toks_[pos_ + 2].text == "^"))) {
                std::string num;
                if (is_op("-")) {
                    num = "-";
                    ++pos_;
                }
                num += toks_[pos_++].text;
                if (is_op("^")) {  // a^2^3 = a^(2^3): fall back to the general form
                    ++pos_;
                    const std::string rest = unary();
                    return "d_pow(" + a + ", d_pow(D::c(" + num + "), " + rest + "))";
                }
                return "d_powc(" + a + ", " + num + ")";
            }
            return "d_pow(" + a + ", " + unary() + ")";
        }
        return a;
    }
    std::string primary() {
        const Token t = toks_[pos_];
        if (t.kind == Token::NUM) {
            ++pos_;
            return "D::c(" + t.text + ")";
        }
        if (t.kind == Token::NAME) {
            ++pos_;
            if (is_op("(")) {
                auto tab = tabs_.find(t.text);
                if (tab != tabs_.end()) {
                    ++pos_;
                    std::string call = tab->second.first + "(" + sum();
                    int nargs = 1;
                    while (is_op(",")) {
                        ++pos_;
                        call += ", " + sum();
                        ++nargs;
                    }
                    expect(")");
                    if (nargs != tab->second.second)
                        fail("tabulated function '" + t.text + "' takes " + std::to_string(tab->second.second) +
                             " argument(s)");
                    return call + ")";
                }
                auto it = functions().find(t.text);
                if (it == functions().end()) fail("unknown function '" + t.text + "'");
                ++pos_;
                std::vector<std::string> args;
                if (!is_op(")")) {
                    args.push_back(sum());
                    while (is_op(",")) {
                        ++pos_;
                        args.push_back(sum());
                    }
                }
                expect(")");
                if ((int)args.size() != it->second.first)
                    fail("function '" + t.text + "' takes " + std::to_string(it->second.first) + " argument(s)");
                std::string call = std::string(it->second.second) + "(";
                for (size_t k = 0; k < args.size(); ++k) call += (k ? ", " : "") + args[k];
                return call + ")";
            }
            auto v = vars_.find(t.text);
            if (v == vars_.end()) fail("unknown variable '" + t.text + "'");
            return v->second;
        }
        if (is_op("(")) {
            ++pos_;
            std::string e = sum();
            expect(")");
            return e;
        }
        fail(t.kind == Token::END ? std::string("unexpected end") : "unexpected '" + t.text + "'");
    }

    std::string src_;
    std::vector<Token> toks_;
    std::map<std::string, std::string> vars_;
    std::map<std::string, std::pair<std::string, int>> tabs_;
    size_t pos_ = 0;
};

bool valid_name(const std::string& n) {
    if (n.empty() || !(std::isalpha((unsigned char)n[0]) || n[0] == '_')) return false;
    for (char c : n)
        if (!(std::isalnum((unsigned char)c) || c == '_')) return false;
    return functions().find(n) == functions().end();
}

// The dual-number type and its functions (device source, prepended to the kernels).
const char* kDualSource = R"(
#pragma clang fp contract(off)
#define GBI_M1(fn) \
    __device__ inline float m_##fn(float x) { return fn##f(x); } \
    __device__ inline double m_##fn(double x) { return fn(x); }
GBI_M1(sqrt) GBI_M1(exp) GBI_M1(log) GBI_M1(sin) GBI_M1(cos) GBI_M1(tan) GBI_M1(asin) GBI_M1(acos)
GBI_M1(atan) GBI_M1(sinh) GBI_M1(cosh) GBI_M1(tanh) GBI_M1(erf) GBI_M1(erfc) GBI_M1(fabs)
__device__ inline float m_pow(float a, float b) { return powf(a, b); }
__device__ inline double m_pow(double a, double b) { return pow(a, b); }
template <typename T, int M> struct Du {
    T v;
    T d[M];
    static __device__ inline Du c(T x) { Du r; r.v = x; for (int k = 0; k < M; ++k) r.d[k] = T(0); return r; }
    static __device__ inline Du var(T x, int slot) { Du r = c(x); r.d[slot] = T(1); return r; }
};
#define DU template <typename T, int M> __device__ inline Du<T, M>
DU d_chain(const Du<T, M>& a, T f, T df) {
    Du<T, M> r; r.v = f; for (int k = 0; k < M; ++k) r.d[k] = df * a.d[k]; return r;
}
DU operator+(const Du<T, M>& a, const Du<T, M>& b) {
    Du<T, M> r; r.v = a.v + b.v; for (int k = 0; k < M; ++k) r.d[k] = a.d[k] + b.d[k]; return r;
}
DU operator-(const Du<T, M>& a, const Du<T, M>& b) {
    Du<T, M> r; r.v = a.v - b.v; for (int k = 0; k < M; ++k) r.d[k] = a.d[k] - b.d[k]; return r;
}
DU operator-(const Du<T, M>& a) {
    Du<T, M> r; r.v = -a.v; for (int k = 0; k < M; ++k) r.d[k] = -a.d[k]; return r;
}
DU operator*(const Du<T, M>& a, const Du<T, M>& b) {
    Du<T, M> r; r.v = a.v * b.v; for (int k = 0; k < M; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r;
}
DU operator/(const Du<T, M>& a, const Du<T, M>& b) {
    Du<T, M> r; r.v = a.v / b.v; const T ib = T(1) / b.v;
    for (int k = 0; k < M; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib; return r;
}
DU d_sqrt(const Du<T, M>& a) { const T f = m_sqrt(a.v); return d_chain(a, f, T(0.5) / f); }
DU d_exp(const Du<T, M>& a) { const T f = m_exp(a.v); return d_chain(a, f, f); }
DU d_log(const Du<T, M>& a) { return d_chain(a, m_log(a.v), T(1) / a.v); }
DU d_sin(const Du<T, M>& a) { return d_chain(a, m_sin(a.v), m_cos(a.v)); }
DU d_cos(const Du<T, M>& a) { return d_chain(a, m_cos(a.v), -m_sin(a.v)); }
DU d_tan(const Du<T, M>& a) { const T f = m_tan(a.v); return d_chain(a, f, T(1) + f * f); }
DU d_asin(const Du<T, M>& a) { return d_chain(a, m_asin(a.v), T(1) / m_sqrt(T(1) - a.v * a.v)); }
DU d_acos(const Du<T, M>& a) { return d_chain(a, m_acos(a.v), T(-1) / m_sqrt(T(1) - a.v * a.v)); }
DU d_atan(const Du<T, M>& a) { return d_chain(a, m_atan(a.v), T(1) / (T(1) + a.v * a.v)); }
DU d_sinh(const Du<T, M>& a) { return d_chain(a, m_sinh(a.v), m_cosh(a.v)); }
DU d_cosh(const Du<T, M>& a) { return d_chain(a, m_cosh(a.v), m_sinh(a.v)); }
DU d_tanh(const Du<T, M>& a) { const T f = m_tanh(a.v); return d_chain(a, f, T(1) - f * f); }
DU d_erf(const Du<T, M>& a) { return d_chain(a, m_erf(a.v), T(1.1283791670955126) * m_exp(-a.v * a.v)); }
DU d_erfc(const Du<T, M>& a) { return d_chain(a, m_erfc(a.v), T(-1.1283791670955126) * m_exp(-a.v * a.v)); }
DU d_abs(const Du<T, M>& a) { return d_chain(a, m_fabs(a.v), a.v < T(0) ? T(-1) : T(1)); }
DU d_step(const Du<T, M>& a) { return Du<T, M>::c(a.v >= T(0) ? T(1) : T(0)); }
DU d_delta(const Du<T, M>& a) { return Du<T, M>::c(a.v == T(0) ? T(1) : T(0)); }
DU d_square(const Du<T, M>& a) { return d_chain(a, a.v * a.v, T(2) * a.v); }
DU d_cube(const Du<T, M>& a) { return d_chain(a, a.v * a.v * a.v, T(3) * a.v * a.v); }
DU d_recip(const Du<T, M>& a) { return d_chain(a, T(1) / a.v, T(-1) / (a.v * a.v)); }
DU d_min(const Du<T, M>& a, const Du<T, M>& b) { return b.v < a.v ? b : a; }
DU d_max(const Du<T, M>& a, const Du<T, M>& b) { return b.v > a.v ? b : a; }
DU d_select(const Du<T, M>& x, const Du<T, M>& a, const Du<T, M>& b) { return x.v != T(0) ? a : b; }
DU d_powc(const Du<T, M>& a, double e) {
    const T te = T(e);
    return d_chain(a, m_pow(a.v, te), te == T(0) ? T(0) : te * m_pow(a.v, te - T(1)));
}
DU d_pow(const Du<T, M>& a, const Du<T, M>& b) {
    const T f = m_pow(a.v, b.v);
    const T da = b.v == T(0) ? T(0) : b.v * m_pow(a.v, b.v - T(1));
    Du<T, M> r; r.v = f;
    for (int k = 0; k < M; ++k) r.d[k] = da * a.d[k] + (b.d[k] != T(0) ? f * m_log(a.v) * b.d[k] : T(0));
    return r;
}
// cubic spline through y at lo + k h (second derivatives y2), evaluated in FP64: natural and
// zero outside [lo, hi], or periodic with period hi - lo
DU d_tabc(const Du<T, M>& a, const double* y, const double* y2, int n, double lo, double hi, int periodic) {
    double t = (double)a.v;
    if (periodic) t -= (hi - lo) * floor((t - lo) / (hi - lo));
    else if (!(t >= lo && t <= hi)) return Du<T, M>::c(T(0));
    const double h = (hi - lo) / (n - 1);
    int k = (int)floor((t - lo) / h);
    k = k < 0 ? 0 : (k > n - 2 ? n - 2 : k);
    const double A = (lo + (k + 1) * h - t) / h, B = 1.0 - A;
    const double f = A * y[k] + B * y[k + 1] + ((A * A * A - A) * y2[k] + (B * B * B - B) * y2[k + 1]) * (h * h / 6.0);
    const double df = (y[k + 1] - y[k]) / h + ((1.0 - 3.0 * A * A) * y2[k] + (3.0 * B * B - 1.0) * y2[k + 1]) * (h / 6.0);
    return d_chain(a, T(f), T(df));
}
// y[round(x) + nx (round(y) + ny round(z))]; NaN when a rounded index leaves its size (the
// evaluation then reports a non-finite result); zero derivatives
__device__ inline bool tab_index(double v, int n, int& k) {
    const double t = round(v);
    k = (int)t;
    return t >= 0.0 && t < (double)n;
}
DU d_tabd(const Du<T, M>& a, const double* y, int n) {
    int i;
    return Du<T, M>::c(tab_index((double)a.v, n, i) ? T(y[i]) : T(__builtin_nan("")));
}
DU d_tabd2(const Du<T, M>& a, const Du<T, M>& b, const double* y, int nx, int ny) {
    int i, j;
    const bool ok = tab_index((double)a.v, nx, i) & tab_index((double)b.v, ny, j);
    return Du<T, M>::c(ok ? T(y[i + nx * j]) : T(__builtin_nan("")));
}
DU d_tabd3(const Du<T, M>& a, const Du<T, M>& b, const Du<T, M>& c, const double* y, int nx, int ny, int nz) {
    int i, j, k;
    const bool ok = tab_index((double)a.v, nx, i) & tab_index((double)b.v, ny, j) & tab_index((double)c.v, nz, k);
    return Du<T, M>::c(ok ? T(y[i + nx * (j + ny * k)]) : T(__builtin_nan("")));
}
// Continuous2D/3D: the knot interval of v on one axis (wrapped when periodic; false outside
// the range otherwise), its width h and the local coordinate t in [0, 1]
__device__ inline bool tab_cell(double v, double lo, double hi, int n, int periodic, int& k, double& t, double& h) {
    if (periodic) v -= (hi - lo) * floor((v - lo) / (hi - lo));
    else if (!(v >= lo && v <= hi)) return false;
    h = (hi - lo) / (n - 1);
    k = (int)floor((v - lo) / h);
    k = k < 0 ? 0 : (k > n - 2 ? n - 2 : k);
    t = (v - (lo + k * h)) / h;
    return true;
}
// cubic Hermite weights on [0, 1]: w[0], w[1] of the values at 0 and 1, w[2], w[3] of the
// slopes at 0 and 1 (per unit t); dw their derivatives in t
__device__ inline void hermite_w(double t, double* w, double* dw) {
    const double t2 = t * t, t3 = t2 * t;
    w[0] = 2.0 * t3 - 3.0 * t2 + 1.0;
    w[1] = 3.0 * t2 - 2.0 * t3;
    w[2] = t3 - 2.0 * t2 + t;
    w[3] = t3 - t2;
    dw[0] = 6.0 * t2 - 6.0 * t;
    dw[1] = 6.0 * t - 6.0 * t2;
    dw[2] = 3.0 * t2 - 4.0 * t + 1.0;
    dw[3] = 3.0 * t2 - 2.0 * t;
}
// bicubic patch from the knot tables f, f_x, f_y, f_xy
DU d_tabc2(const Du<T, M>& a, const Du<T, M>& b, const double* f, const double* fx, const double* fy, const double* fxy,
           int nx, int ny, double x0, double x1, double y0, double y1, int periodic) {
    int i, j;
    double t, u, hx, hy;
    if (!tab_cell((double)a.v, x0, x1, nx, periodic, i, t, hx) || !tab_cell((double)b.v, y0, y1, ny, periodic, j, u, hy))
        return Du<T, M>::c(T(0));
    double wx[4], dwx[4], wy[4], dwy[4];
    hermite_w(t, wx, dwx);
    hermite_w(u, wy, dwy);
    double v = 0.0, gx = 0.0, gy = 0.0;
    for (int cb = 0; cb < 2; ++cb)
        for (int ca = 0; ca < 2; ++ca) {
            const int e = (i + ca) + nx * (j + cb);
            const double c00 = f[e], c10 = fx[e] * hx, c01 = fy[e] * hy, c11 = fxy[e] * hx * hy;
            const double X0 = wx[ca], X1 = wx[2 + ca], Y0 = wy[cb], Y1 = wy[2 + cb];
            v += c00 * X0 * Y0 + c10 * X1 * Y0 + c01 * X0 * Y1 + c11 * X1 * Y1;
            gx += c00 * dwx[ca] * Y0 + c10 * dwx[2 + ca] * Y0 + c01 * dwx[ca] * Y1 + c11 * dwx[2 + ca] * Y1;
            gy += c00 * X0 * dwy[cb] + c10 * X1 * dwy[cb] + c01 * X0 * dwy[2 + cb] + c11 * X1 * dwy[2 + cb];
        }
    gx /= hx;
    gy /= hy;
    Du<T, M> r;
    r.v = T(v);
    for (int k = 0; k < M; ++k) r.d[k] = T(gx) * a.d[k] + T(gy) * b.d[k];
    return r;
}
// tricubic patch from the knot tables g[0..7] = f, f_x, f_y, f_z, f_xy, f_xz, f_yz, f_xyz
DU d_tabc3(const Du<T, M>& a, const Du<T, M>& b, const Du<T, M>& c, const double* f, const double* fx,
           const double* fy, const double* fz, const double* fxy, const double* fxz, const double* fyz,
           const double* fxyz, int nx, int ny, int nz, double x0, double x1, double y0, double y1, double z0,
           double z1, int periodic) {
    int i, j, l;
    double t, u, s, hx, hy, hz;
    if (!tab_cell((double)a.v, x0, x1, nx, periodic, i, t, hx) || !tab_cell((double)b.v, y0, y1, ny, periodic, j, u, hy) ||
        !tab_cell((double)c.v, z0, z1, nz, periodic, l, s, hz))
        return Du<T, M>::c(T(0));
    double wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
    hermite_w(t, wx, dwx);
    hermite_w(u, wy, dwy);
    hermite_w(s, wz, dwz);
    const double* g[8] = {f, fx, fy, fz, fxy, fxz, fyz, fxyz};
    const int px[8] = {0, 1, 0, 0, 1, 1, 0, 1}, py[8] = {0, 0, 1, 0, 1, 0, 1, 1}, pz[8] = {0, 0, 0, 1, 0, 1, 1, 1};
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int cc = 0; cc < 2; ++cc)
        for (int cb = 0; cb < 2; ++cb)
            for (int ca = 0; ca < 2; ++ca) {
                const int e = (i + ca) + nx * ((j + cb) + ny * (l + cc));
                for (int m = 0; m < 8; ++m) {
                    const double k = g[m][e] * (px[m] ? hx : 1.0) * (py[m] ? hy : 1.0) * (pz[m] ? hz : 1.0);
                    const double X = wx[2 * px[m] + ca], Y = wy[2 * py[m] + cb], Z = wz[2 * pz[m] + cc];
                    v += k * X * Y * Z;
                    gx += k * dwx[2 * px[m] + ca] * Y * Z;
                    gy += k * X * dwy[2 * py[m] + cb] * Z;
                    gz += k * X * Y * dwz[2 * pz[m] + cc];
                }
            }
    gx /= hx;
    gy /= hy;
    gz /= hz;
    Du<T, M> r;
    r.v = T(v);
    for (int k = 0; k < M; ++k) r.d[k] = T(gx) * a.d[k] + T(gy) * b.d[k] + T(gz) * c.d[k];
    return r;
}
)";

void check_hiprtc(hiprtcResult r, const char* what) {
    if (r != HIPRTC_SUCCESS) throw std::runtime_error(std::string(what) + ": " + hiprtcGetErrorString(r));
}

}  // namespace

CustomProgram::~CustomProgram() {
    if (module) (void)hipModuleUnload(module);
}

std::string translate_expression(const std::string& expr, const CustomSpec& spec, int context, int index) {
    std::map<std::string, std::string> vars;
    std::map<std::string, std::pair<std::string, int>> tabs;
    for (size_t k = 0; k < spec.functions.size(); ++k)
        tabs[spec.functions[k].name] = {"d_tab" + std::to_string(k), tab_arity(spec.functions[k].kind)};
    for (size_t k = 0; k < spec.globals.size(); ++k) vars[spec.globals[k]] = "G" + std::to_string(k);
    if (context == 2) {
        vars["r"] = "r_";
        for (size_t k = 0; k < spec.params.size(); ++k) {
            vars[spec.params[k] + "1"] = "P1_" + std::to_string(k);
            vars[spec.params[k] + "2"] = "P2_" + std::to_string(k);
        }
        for (size_t k = 0; k < spec.value_names.size(); ++k) {
            vars[spec.value_names[k] + "1"] = "V1_" + std::to_string(k);
            vars[spec.value_names[k] + "2"] = "V2_" + std::to_string(k);
        }
    } else {
        vars["x"] = "X_";
        vars["y"] = "Y_";
        vars["z"] = "Z_";
        for (size_t k = 0; k < spec.params.size(); ++k) vars[spec.params[k]] = "P_" + std::to_string(k);
        for (size_t k = 0; k < spec.integrals.size(); ++k) vars[spec.integrals[k]] = "I_" + std::to_string(k);
        const int nv = context == 0 ? index : (int)spec.value_names.size();
        for (int k = 0; k < nv; ++k) vars[spec.value_names[k]] = "V_" + std::to_string(k);
    }
    return Parser(expr, vars, tabs).parse();
}

std::vector<double> spline_second_derivatives(const std::vector<double>& y, double h, bool periodic) {
    const int n = (int)y.size();
    if (n < 2) throw std::runtime_error("a tabulated function needs at least 2 values");
    if (!(h > 0.0) || !std::isfinite(h)) throw std::runtime_error("a tabulated function needs min < max");
    std::vector<double> M(n, 0.0);
    if (!periodic) {
        // M_0 = M_{n-1} = 0; M_{i-1} + 4 M_i + M_{i+1} = 6 (y_{i+1} - 2 y_i + y_{i-1}) / h^2
        if (n < 3) return M;
        const int m = n - 2;
        std::vector<double> c(m), d(m);
        for (int i = 0; i < m; ++i) {
            const double rhs = 6.0 * (y[i + 2] - 2.0 * y[i + 1] + y[i]) / (h * h);
            const double den = 4.0 - (i > 0 ? c[i - 1] : 0.0);
            c[i] = 1.0 / den;
            d[i] = (rhs - (i > 0 ? d[i - 1] : 0.0)) / den;
        }
        for (int i = m - 1; i >= 0; --i) M[i + 1] = d[i] - (i + 1 < m ? c[i] * M[i + 2] : 0.0);
        return M;
    }
    if (y.front() != y.back())
        throw std::runtime_error("a periodic tabulated function needs equal first and last values");
    // the same equations with indices modulo m = n - 1 (M_{n-1} = M_0): a cyclic tridiagonal
    // system, solved by Sherman-Morrison from two ordinary tridiagonal solves
    const int m = n - 1;
    auto rhs = [&](int i) {
        const double ym = y[(i - 1 + m) % m], y0 = y[i], yp = y[(i + 1) % m];
        return 6.0 * (yp - 2.0 * y0 + ym) / (h * h);
    };
    if (m <= 2) {  // 1 or 2 unknowns: constant or two-point tables have no curvature to spread
        if (m == 1) M[0] = rhs(0) / 6.0;  // M_0 (1 + 4 + 1) = rhs
        if (m == 2) {
            // [4 2; 2 4] M = rhs
            const double r0 = rhs(0), r1 = rhs(1);
            M[0] = (4.0 * r0 - 2.0 * r1) / 12.0;
            M[1] = (4.0 * r1 - 2.0 * r0) / 12.0;
        }
        M[m] = M[0];
        return M;
    }
    // A = T + u v^T with T tridiagonal (diagonal 4, first entry 4 - g, last 4 - 1/g), u = (g, 0, .., 1),
    // v = (1, 0, .., 1/g), g = -4
    const double g = -4.0;
    std::vector<double> diag(m, 4.0), b(m), u(m, 0.0);
    diag[0] -= g;
    diag[m - 1] -= 1.0 / g;
    for (int i = 0; i < m; ++i) b[i] = rhs(i);
    u[0] = g;
    u[m - 1] = 1.0;
    auto solve = [&](std::vector<double> r) {
        std::vector<double> c(m), x(m);
        double den = diag[0];
        c[0] = 1.0 / den;
        r[0] /= den;
        for (int i = 1; i < m; ++i) {
            den = diag[i] - c[i - 1];
            c[i] = 1.0 / den;
            r[i] = (r[i] - r[i - 1]) / den;
        }
        x[m - 1] = r[m - 1];
        for (int i = m - 2; i >= 0; --i) x[i] = r[i] - c[i] * x[i + 1];
        return x;
    };
    const std::vector<double> xb = solve(b), zu = solve(u);
    const double fact = (xb[0] + xb[m - 1] / g) / (1.0 + zu[0] + zu[m - 1] / g);
    for (int i = 0; i < m; ++i) M[i] = xb[i] - fact * zu[i];
    M[m] = M[0];
    return M;
}

std::vector<double> spline_knot_derivatives(const std::vector<double>& y, double h, bool periodic) {
    const std::vector<double> M = spline_second_derivatives(y, h, periodic);
    const int n = (int)y.size();
    std::vector<double> d(n);
    // S'(x_i) from the piece to the right of knot i (the last knot: the piece to its left)
    for (int i = 0; i + 1 < n; ++i) d[i] = (y[i + 1] - y[i]) / h - h * (2.0 * M[i] + M[i + 1]) / 6.0;
    d[n - 1] = (y[n - 1] - y[n - 2]) / h + h * (M[n - 2] + 2.0 * M[n - 1]) / 6.0;
    return d;
}

std::vector<std::vector<double>> hermite_tables(const TabFunction& f) {
    const bool three = f.kind == TAB_CONTINUOUS_3D;
    const int n[3] = {f.dims[0], f.dims[1], three ? f.dims[2] : 1};
    const double lo[3] = {f.min, f.ymin, f.zmin}, hi[3] = {f.max, f.ymax, f.zmax};
    double h[3];
    for (int a = 0; a < (three ? 3 : 2); ++a) {
        if (n[a] < 2) throw std::runtime_error("a continuous tabulated function needs at least 2 values per axis");
        h[a] = (hi[a] - lo[a]) / (n[a] - 1);
    }
    // derivative along `axis` at the knots of every line of the table
    auto along = [&](const std::vector<double>& in, int axis) {
        std::vector<double> out(in.size()), line(n[axis]);
        const int stride = axis == 0 ? 1 : (axis == 1 ? n[0] : n[0] * n[1]);
        for (int z = 0; z < n[2]; ++z)
            for (int y = 0; y < n[1]; ++y)
                for (int x = 0; x < n[0]; ++x) {
                    const int first = x + n[0] * (y + n[1] * z);
                    if ((axis == 0 && x) || (axis == 1 && y) || (axis == 2 && z)) continue;  // one pass per line
                    for (int k = 0; k < n[axis]; ++k) line[k] = in[first + k * stride];
                    const std::vector<double> d = spline_knot_derivatives(line, h[axis], f.periodic);
                    for (int k = 0; k < n[axis]; ++k) out[first + k * stride] = d[k];
                }
        return out;
    };
    const std::vector<double>& v = f.values;
    const std::vector<double> fx = along(v, 0), fy = along(v, 1), fxy = along(fx, 1);
    if (!three) return {v, fx, fy, fxy};
    const std::vector<double> fz = along(v, 2), fxz = along(fx, 2), fyz = along(fy, 2), fxyz = along(fxy, 2);
    return {v, fx, fy, fz, fxy, fxz, fyz, fxyz};
}

void generate_custom_source(const CustomSpec& spec, CustomProgram& out) {
    const int NP = (int)spec.params.size(), NG = (int)spec.globals.size(), NI = (int)spec.integrals.size();
    const int NV = (int)spec.value_names.size();
    if (NI < 1 || NI > kMaxCustomIntegrals) throw std::runtime_error("one or two integrals are supported");
    if (NV > kMaxCustomValues) throw std::runtime_error("at most 4 computed values are supported");
    if (spec.value_exprs.size() != spec.value_names.size() || spec.term_exprs.size() != spec.term_types.size())
        throw std::runtime_error("inconsistent expression lists");
    std::map<std::string, int> seen;
    auto name_ok = [&](const std::string& n, const char* what) {
        if (!valid_name(n)) throw std::runtime_error(std::string("invalid ") + what + " name '" + n + "'");
        if (n == "r") throw std::runtime_error("'r' is reserved for the pair distance");
        if (n == "x" || n == "y" || n == "z") throw std::runtime_error("'" + n + "' is reserved for the atom's coordinate");
        if (seen.count(n)) throw std::runtime_error("name '" + n + "' defined twice");
        seen[n] = 1;
    };
    for (const auto& f : spec.functions) {
        name_ok(f.name, "tabulated function");
        if (f.kind < TAB_CONTINUOUS_1D || f.kind > TAB_CONTINUOUS_3D)
            throw std::runtime_error("tabulated function '" + f.name + "': unknown kind");
        if (f.values.empty()) throw std::runtime_error("tabulated function '" + f.name + "' has no values");
        const int ar = tab_arity(f.kind);
        if (ar > 1) {
            const long long want = (long long)f.dims[0] * f.dims[1] * (ar == 3 ? f.dims[2] : 1);
            const int least = f.kind == TAB_CONTINUOUS_2D || f.kind == TAB_CONTINUOUS_3D ? 2 : 1;
            if (f.dims[0] < least || f.dims[1] < least || (ar == 3 && f.dims[2] < least) ||
                want != (long long)f.values.size())
                throw std::runtime_error("tabulated function '" + f.name + "': the sizes do not match the values");
        }
        if (f.kind == TAB_CONTINUOUS_2D || f.kind == TAB_CONTINUOUS_3D) {
            if (!(f.min < f.max) || !(f.ymin < f.ymax) || (ar == 3 && !(f.zmin < f.zmax)))
                throw std::runtime_error("tabulated function '" + f.name + "' needs min < max on every axis");
        }
        for (double v : f.values)
            if (!std::isfinite(v)) throw std::runtime_error("tabulated function '" + f.name + "' has a non-finite value");
    }
    for (const auto& n : spec.params) name_ok(n, "parameter");
    for (const auto& n : spec.globals) name_ok(n, "global parameter");
    for (const auto& n : spec.integrals) name_ok(n, "integral");
    for (const auto& n : spec.value_names) name_ok(n, "computed value");
    for (int t : spec.term_types)
        if (t < TERM_SINGLE || t > TERM_PAIR_NO_EXCL) throw std::runtime_error("unknown energy term type");
    const int ND = (int)spec.deriv_globals.size();
    if (ND > kMaxCustomDerivs) throw std::runtime_error("at most 4 energy parameter derivatives are supported");
    std::vector<int> deriv_slot(NG, -1);  // global k -> its derivative slot
    for (int m = 0; m < ND; ++m) {
        const int g = spec.deriv_globals[m];
        if (g < 0 || g >= NG || deriv_slot[g] >= 0)
            throw std::runtime_error("an energy parameter derivative must name a global parameter, once");
        deriv_slot[g] = m;
    }

    std::ostringstream s;
    s << kDualSource;
    // the tabulated functions' tables (FP64 literals, exact to the last bit) and callers
    s << std::setprecision(17);
    for (size_t k = 0; k < spec.functions.size(); ++k) {
        const TabFunction& f = spec.functions[k];
        const int n = (int)f.values.size();
        auto table = [&](const char* tag, const std::vector<double>& v) {
            s << "__device__ const double TAB" << k << tag << "[" << n << "] = {";
            for (int i = 0; i < n; ++i) s << (i ? ", " : "") << v[i];
            s << "};\n";
        };
        table("_Y", f.values);
        if (f.kind == TAB_CONTINUOUS_1D) {
            const std::vector<double> y2 =
                spline_second_derivatives(f.values, (f.max - f.min) / (n - 1), f.periodic);
            table("_Y2", y2);
            s << "DU d_tab" << k << "(const Du<T, M>& a) { return d_tabc(a, TAB" << k << "_Y, TAB" << k << "_Y2, " << n
              << ", " << f.min << ", " << f.max << ", " << (f.periodic ? 1 : 0) << "); }\n";
        } else if (f.kind == TAB_DISCRETE_1D) {
            s << "DU d_tab" << k << "(const Du<T, M>& a) { return d_tabd(a, TAB" << k << "_Y, " << n << "); }\n";
        } else if (f.kind == TAB_DISCRETE_2D) {
            s << "DU d_tab" << k << "(const Du<T, M>& a, const Du<T, M>& b) { return d_tabd2(a, b, TAB" << k << "_Y, "
              << f.dims[0] << ", " << f.dims[1] << "); }\n";
        } else if (f.kind == TAB_DISCRETE_3D) {
            s << "DU d_tab" << k << "(const Du<T, M>& a, const Du<T, M>& b, const Du<T, M>& c) { return d_tabd3(a, b, c, TAB"
              << k << "_Y, " << f.dims[0] << ", " << f.dims[1] << ", " << f.dims[2] << "); }\n";
        } else {
            // knot tables of the bicubic / tricubic patch (the values are TABk_Y)
            const std::vector<std::vector<double>> ht = hermite_tables(f);
            static const char* tags[8] = {"_Y", "_X", "_YD", "_Z", "_XY", "_XZ", "_YZ", "_XYZ"};
            const bool three = f.kind == TAB_CONTINUOUS_3D;
            // 2D order f, f_x, f_y, f_xy; 3D order f, f_x, f_y, f_z, f_xy, f_xz, f_yz, f_xyz
            const int tag_of[8] = {0, 1, 2, three ? 3 : 4, 4, 5, 6, 7};
            for (size_t m = 1; m < ht.size(); ++m) table(tags[tag_of[m]], ht[m]);
            std::ostringstream arg;
            arg << std::setprecision(17);
            for (size_t m = 0; m < ht.size(); ++m) arg << "TAB" << k << tags[tag_of[m]] << ", ";
            if (!three) {
                s << "DU d_tab" << k << "(const Du<T, M>& a, const Du<T, M>& b) { return d_tabc2(a, b, " << arg.str()
                  << f.dims[0] << ", " << f.dims[1] << ", " << f.min << ", " << f.max << ", " << f.ymin << ", "
                  << f.ymax << ", " << (f.periodic ? 1 : 0) << "); }\n";
            } else {
                s << "DU d_tab" << k << "(const Du<T, M>& a, const Du<T, M>& b, const Du<T, M>& c) { return d_tabc3(a, b, c, "
                  << arg.str() << f.dims[0] << ", " << f.dims[1] << ", " << f.dims[2] << ", " << f.min << ", "
                  << f.max << ", " << f.ymin << ", " << f.ymax << ", " << f.zmin << ", " << f.zmax << ", "
                  << (f.periodic ? 1 : 0) << "); }\n";
            }
        }
    }
    s << std::setprecision(6);
    s << "#define NP " << NP << "\n#define NG " << (NG > 0 ? NG : 1) << "\n#define NV " << NV << "\n#define ND "
      << ND << "\n#define NDD " << (ND > 0 ? ND : 1) << "\n";
    const int NPL = (4 + NV + ND) + (3 + NV);  // row and column planes of a pair tile

    // ---- per atom: computed values and single-particle terms, duals over the integrals and
    // the globals whose energy derivative is asked for ----
    // dual slots: 0, 1 the integrals, 2 .. 2 + ND - 1 the globals of the energy derivatives,
    // XS .. XS + 2 the atom's x, y, z
    s << "#define XS (2 + ND)\n";
    s << "extern \"C\" __global__ void cg_values(int N, const double* __restrict__ prm, const double* __restrict__ "
         "integ, const double* __restrict__ glob, double* __restrict__ vals, double* __restrict__ dvdi, double* "
         "__restrict__ esp, double* __restrict__ despdi, double* __restrict__ dvdth, double* __restrict__ despdth, "
         "const double* __restrict__ pos, double* __restrict__ dvdx, double* __restrict__ despdx) {\n"
         "    const int i = blockIdx.x * blockDim.x + threadIdx.x;\n    if (i >= N) return;\n"
         "    typedef Du<double, XS + 3> D;\n"
         "    const D X_ = D::var(pos[3 * i], XS), Y_ = D::var(pos[3 * i + 1], XS + 1), Z_ = D::var(pos[3 * i + 2], XS + 2);\n";
    for (int k = 0; k < NG; ++k) {
        if (deriv_slot[k] >= 0)
            s << "    const D G" << k << " = D::var(glob[" << k << "], " << 2 + deriv_slot[k] << ");\n";
        else
            s << "    const D G" << k << " = D::c(glob[" << k << "]);\n";
    }
    for (int k = 0; k < NP; ++k) s << "    const D P_" << k << " = D::c(prm[(size_t)i * NP + " << k << "]);\n";
    for (int k = 0; k < NI; ++k) s << "    const D I_" << k << " = D::var(integ[" << k << " * N + i], " << k << ");\n";
    for (int k = 0; k < NV; ++k) {
        s << "    const D V_" << k << " = " << translate_expression(spec.value_exprs[k], spec, 0, k) << ";\n";
        s << "    vals[" << k << " * N + i] = V_" << k << ".v;\n";
        s << "    dvdi[(" << k << " * 2 + 0) * N + i] = V_" << k << ".d[0];\n";
        s << "    dvdi[(" << k << " * 2 + 1) * N + i] = V_" << k << ".d[1];\n";
        s << "    for (int m = 0; m < ND; ++m) dvdth[(" << k << " * NDD + m) * N + i] = V_" << k << ".d[2 + m];\n";
        s << "    for (int d = 0; d < 3; ++d) dvdx[(" << k << " * 3 + d) * N + i] = V_" << k << ".d[XS + d];\n";
    }
    s << "    D E_ = D::c(0.0);\n";
    for (size_t t = 0; t < spec.term_exprs.size(); ++t)
        if (spec.term_types[t] == TERM_SINGLE)
            s << "    E_ = E_ + (" << translate_expression(spec.term_exprs[t], spec, 1, 0) << ");\n";
    s << "    esp[i] = E_.v;\n    despdi[i] = E_.d[0];\n    despdi[N + i] = E_.d[1];\n"
         "    for (int m = 0; m < ND; ++m) despdth[m * N + i] = E_.d[2 + m];\n"
         "    for (int d = 0; d < 3; ++d) despdx[d * N + i] = E_.d[XS + d];\n}\n";

    // ---- pairs: 64 x 64 tiles (row block I, column block J >= I), each pair once with atom1
    // the row atom (atom1 < atom2 as the reference orders a pair, CharmmReferenceGBMV.cpp:
    // 279-286).  The pair difference, r^2, the minimum image and the cutoff test are FP64;
    // r and the expressions (FP32 duals) are FP32, as are a wave's partial sums over its
    // CG_STEPS column steps (the tile's 64 steps are split over CG_WAVES waves; within a wave
    // the column sums rotate through the lanes by DPP).  The waves' row and column partials
    // are added in FP64 in LDS in wave order, and cg_chain adds the tiles' partials per atom
    // in FP64 in a fixed order ----
    std::string pair_excl, pair_noexcl;
    for (size_t t = 0; t < spec.term_exprs.size(); ++t) {
        if (spec.term_types[t] == TERM_SINGLE) continue;
        const std::string e = "            E_ = E_ + (" + translate_expression(spec.term_exprs[t], spec, 2, 0) + ");\n";
        (spec.term_types[t] == TERM_PAIR ? pair_excl : pair_noexcl) += e;
        out.any_pair = true;
    }
    s << "#define NPR (4 + NV + ND)\n#define NPC (3 + NV)\n#define CG_WAVES " << kCustomPairWaves << "\n"
         "#define CG_STEPS (64 / CG_WAVES)\n"
         "__device__ inline float rot_down(float v) {\n"
         "    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));\n}\n"
         "__device__ inline long long tile_index(int I, int J, int nb) {\n"
         "    return (long long)I * nb - (long long)I * (I - 1) / 2 + (J - I);\n}\n";
    s << "extern \"C\" __global__ void __launch_bounds__(64 * CG_WAVES) cg_pairs(int N, int nb, const double* __restrict__ pos, "
         "const double* __restrict__ prm, const double* __restrict__ vals, const double* __restrict__ glob, int method, "
         "double cut2, const double* __restrict__ box, const int* __restrict__ ex_ptr, const int* __restrict__ ex_idx, "
         "double* __restrict__ part) {\n"
         "    const int J = blockIdx.x, I = blockIdx.y;\n    if (J < I) return;\n"
         "    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;\n"
         "    __shared__ double s_x[64], s_y[64], s_z[64];\n"
         "    __shared__ double s_red[CG_WAVES][NPR + NPC][64];\n"
         "    __shared__ float s_p[NP][64];\n"
         "    __shared__ float s_v[NV > 0 ? NV : 1][64];\n"
         "    const int i = 64 * I + lane, jl = 64 * J + lane;\n"
         "    const bool iv = i < N, jv = jl < N;\n"
         "    if (w == 0) {\n"
         "        s_x[lane] = jv ? pos[3 * jl] : 0.0;\n        s_y[lane] = jv ? pos[3 * jl + 1] : 0.0;\n"
         "        s_z[lane] = jv ? pos[3 * jl + 2] : 0.0;\n"
         "        for (int k = 0; k < NP; ++k) s_p[k][lane] = jv ? (float)prm[(size_t)jl * NP + k] : 0.0f;\n"
         "        for (int k = 0; k < NV; ++k) s_v[k][lane] = jv ? (float)vals[k * N + jl] : 0.0f;\n"
         "    }\n"
         "    __syncthreads();\n"
         "    typedef Du<float, 1 + 2 * NV + ND> D;\n";
    for (int k = 0; k < NG; ++k) {
        if (deriv_slot[k] >= 0)
            s << "    const D G" << k << " = D::var((float)glob[" << k << "], 1 + 2 * NV + " << deriv_slot[k] << ");\n";
        else
            s << "    const D G" << k << " = D::c(glob[" << k << "]);\n";
    }
    s << "    const double xi = iv ? pos[3 * i] : 0.0, yi = iv ? pos[3 * i + 1] : 0.0, zi = iv ? pos[3 * i + 2] : 0.0;\n";
    for (int k = 0; k < NP; ++k) s << "    const D P1_" << k << " = D::c(iv ? prm[(size_t)i * NP + " << k << "] : 0.0);\n";
    for (int k = 0; k < NV; ++k) s << "    const D V1_" << k << " = D::var(iv ? vals[" << k << " * N + i] : 0.0, " << 1 + k << ");\n";
    s << "    float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;\n"
         "    float dv1[NV > 0 ? NV : 1], cv[NV > 0 ? NV : 1], dth[NDD];\n"
         "    for (int k = 0; k < NV; ++k) dv1[k] = cv[k] = 0.f;\n"
         "    for (int m = 0; m < ND; ++m) dth[m] = 0.f;\n"
         "    const bool has_ex = ex_ptr[N] > 0;\n"
         "    for (int m = CG_STEPS * w; m < CG_STEPS * (w + 1); ++m) {\n"
         "        const int c = (lane + m) & 63;\n"
         "        const int j = 64 * J + c;\n"
         "        bool ok = iv && j < N && j > i;\n"
         "        double dx = 0.0, dy = 0.0, dz = 0.0, r2 = 1.0;\n"
         "        if (ok) {\n"
         "            dx = s_x[c] - xi;\n            dy = s_y[c] - yi;\n            dz = s_z[c] - zi;\n"
         "            if (method == 2) {\n"
         "                dx -= box[0] * floor(dx / box[0] + 0.5);\n"
         "                dy -= box[1] * floor(dy / box[1] + 0.5);\n"
         "                dz -= box[2] * floor(dz / box[2] + 0.5);\n"
         "            }\n"
         "            r2 = dx * dx + dy * dy + dz * dz;\n"
         "            if (method != 0 && !(r2 < cut2)) ok = false;\n"
         "        }\n"
         "        if (ok) {\n"
         "            bool excluded = false;\n"
         "            if (has_ex) {\n"
         "                int lo = ex_ptr[i], hi = ex_ptr[i + 1];\n"
         "                const int end = hi;\n"
         "                while (lo < hi) {\n"
         "                    const int mid = (lo + hi) >> 1;\n"
         "                    if (ex_idx[mid] < j) lo = mid + 1; else hi = mid;\n"
         "                }\n"
         "                excluded = lo < end && ex_idx[lo] == j;\n"
         "            }\n"
         "            const float rr = sqrtf((float)r2);\n"
         "            const D r_ = D::var(rr, 0);\n";
    for (int k = 0; k < NP; ++k) s << "            const D P2_" << k << " = D::c(s_p[" << k << "][c]);\n";
    for (int k = 0; k < NV; ++k) s << "            const D V2_" << k << " = D::var(s_v[" << k << "][c], " << 1 + NV + k << ");\n";
    s << "            D E_ = D::c(0.0);\n" << pair_noexcl << "            if (!excluded) {\n" << pair_excl
      << "            }\n"
         "            e += E_.v;\n"
         "            const float g = E_.d[0] / rr, gx = g * (float)dx, gy = g * (float)dy, gz = g * (float)dz;\n"
         "            fx += gx;\n            fy += gy;\n            fz += gz;\n"
         "            cx -= gx;\n            cy -= gy;\n            cz -= gz;\n"
         "            for (int k = 0; k < NV; ++k) {\n"
         "                dv1[k] += E_.d[1 + k];\n                cv[k] += E_.d[1 + NV + k];\n            }\n"
         "            for (int m = 0; m < ND; ++m) dth[m] += E_.d[1 + 2 * NV + m];\n"
         "        }\n"
         "        cx = rot_down(cx);\n        cy = rot_down(cy);\n        cz = rot_down(cz);\n"
         "        for (int k = 0; k < NV; ++k) cv[k] = rot_down(cv[k]);\n"
         "    }\n"
         "    // after its steps, lane l holds the sums of column (l + CG_STEPS (w + 1)) mod 64\n"
         "    const int cl = (lane + CG_STEPS * (w + 1)) & 63;\n"
         "    s_red[w][0][lane] = fx;\n    s_red[w][1][lane] = fy;\n    s_red[w][2][lane] = fz;\n"
         "    s_red[w][3][lane] = e;\n"
         "    for (int k = 0; k < NV; ++k) s_red[w][4 + k][lane] = dv1[k];\n"
         "    for (int m = 0; m < ND; ++m) s_red[w][4 + NV + m][lane] = dth[m];\n"
         "    s_red[w][NPR][cl] = cx;\n    s_red[w][NPR + 1][cl] = cy;\n    s_red[w][NPR + 2][cl] = cz;\n"
         "    for (int k = 0; k < NV; ++k) s_red[w][NPR + 3 + k][cl] = cv[k];\n"
         "    __syncthreads();\n"
         "    double* o = part + tile_index(I, J, nb) * (NPR + NPC) * 64;\n"
         "    for (int t = threadIdx.x; t < (NPR + NPC) * 64; t += 64 * CG_WAVES) {\n"
         "        double acc = s_red[0][t >> 6][t & 63];\n"
         "        for (int v = 1; v < CG_WAVES; ++v) acc += s_red[v][t >> 6][t & 63];\n"
         "        o[t] = acc;\n    }\n}\n";

    // ---- per 64-atom block I (one workgroup of CG_CHAIN_WAVES waves): the tiles' row and
    // column partials, wave w taking tiles w, w + CG_CHAIN_WAVES, ... and the waves' sums
    // added in wave order (a fixed order), then energy, pair forces, dE/dI per atom ----
    s << "#define CG_CHAIN_WAVES " << kCustomChainWaves << "\n";
    s << "extern \"C\" __global__ void __launch_bounds__(64 * CG_CHAIN_WAVES) cg_chain(int N, int nb, const double* "
         "__restrict__ part, const double* __restrict__ dvdi, const double* __restrict__ esp, const double* __restrict__ "
         "despdi, double* __restrict__ atom_energy, long long* __restrict__ force_fix, double* __restrict__ dedi, "
         "const double* __restrict__ dvdth, const double* __restrict__ despdth, double* __restrict__ pdth, "
         "const double* __restrict__ dvdx, const double* __restrict__ despdx) {\n"
         "    const int I = blockIdx.x, l = threadIdx.x & 63, w = threadIdx.x >> 6;\n"
         "    const int a = 64 * I + l;\n"
         "    __shared__ double s_sum[CG_CHAIN_WAVES][NPR + NPC][64];\n"
         "    double r[NPR], c[NPC];\n"
         "    for (int p = 0; p < NPR; ++p) r[p] = 0.0;\n    for (int p = 0; p < NPC; ++p) c[p] = 0.0;\n"
         "    for (int J = I + w; J < nb; J += CG_CHAIN_WAVES) {\n"
         "        const double* o = part + tile_index(I, J, nb) * (NPR + NPC) * 64;\n"
         "        for (int p = 0; p < NPR; ++p) r[p] += o[p * 64 + l];\n    }\n"
         "    for (int K = w; K <= I; K += CG_CHAIN_WAVES) {\n"
         "        const double* o = part + tile_index(K, I, nb) * (NPR + NPC) * 64;\n"
         "        for (int p = 0; p < NPC; ++p) c[p] += o[(NPR + p) * 64 + l];\n    }\n"
         "    for (int p = 0; p < NPR; ++p) s_sum[w][p][l] = r[p];\n"
         "    for (int p = 0; p < NPC; ++p) s_sum[w][NPR + p][l] = c[p];\n"
         "    __syncthreads();\n"
         "    if (w != 0 || a >= N) return;\n"
         "    for (int v = 1; v < CG_CHAIN_WAVES; ++v) {\n"
         "        for (int p = 0; p < NPR; ++p) r[p] += s_sum[v][p][l];\n"
         "        for (int p = 0; p < NPC; ++p) c[p] += s_sum[v][NPR + p][l];\n    }\n"
         "    atom_energy[a] = esp[a] + r[3];\n"
         "    // pair forces, minus the gradient of the energy with respect to the atom's own\n"
         "    // coordinates through the single-particle terms and through the values (dE/dV_k: the\n"
         "    // pair terms' sums; the single-particle terms' are inside despdx)\n"
         "    for (int d = 0; d < 3; ++d) {\n"
         "        double gx = despdx[d * N + a];\n"
         "        for (int k = 0; k < NV; ++k) gx += (r[4 + k] + c[3 + k]) * dvdx[(k * 3 + d) * N + a];\n"
         "        force_fix[d * N + a] = (long long)llrint((r[d] + c[d] - gx) * 4294967296.0);\n"
         "    }\n"
         "    for (int m = 0; m < 2; ++m) {\n"
         "        double g = despdi[m * N + a];\n"
         "        for (int k = 0; k < NV; ++k) g += (r[4 + k] + c[3 + k]) * dvdi[(k * 2 + m) * N + a];\n"
         "        dedi[m * N + a] = g;\n    }\n"
         "    // dE/dtheta of this atom: its single-particle terms (through the values too), its pair\n"
         "    // terms' explicit derivative (row side: each pair once), and dE/dV dV/dtheta\n"
         "    for (int m = 0; m < ND; ++m) {\n"
         "        double g = despdth[m * N + a] + r[4 + NV + m];\n"
         "        for (int k = 0; k < NV; ++k) g += (r[4 + k] + c[3 + k]) * dvdth[(k * NDD + m) * N + a];\n"
         "        pdth[m * N + a] = g;\n    }\n}\n";
    out.source = s.str();
    out.n_params = NP;
    out.n_globals = NG;
    out.n_values = NV;
    out.n_planes = NPL;
    out.n_derivs = ND;
}

void compile_custom_program(const std::string& arch, CustomProgram& out) {
    hiprtcProgram prog;
    check_hiprtc(hiprtcCreateProgram(&prog, out.source.c_str(), "gbi_custom.hip", 0, nullptr, nullptr),
                 "hiprtcCreateProgram");
    const std::string arch_opt = "--offload-arch=" + arch;
    // packed FP32 stays disabled, as for the library's own kernels
    const char* opts[] = {arch_opt.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult cr = hiprtcCompileProgram(prog, 6, opts);
    if (cr != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw std::runtime_error("generated kernels did not compile: " + log);
    }
    size_t size = 0;
    check_hiprtc(hiprtcGetCodeSize(prog, &size), "hiprtcGetCodeSize");
    std::vector<char> code(size);
    check_hiprtc(hiprtcGetCode(prog, code.data()), "hiprtcGetCode");
    hiprtcDestroyProgram(&prog);
    if (hipModuleLoadData(&out.module, code.data()) != hipSuccess) throw std::runtime_error("hipModuleLoadData failed");
    if (hipModuleGetFunction(&out.values, out.module, "cg_values") != hipSuccess ||
        hipModuleGetFunction(&out.pairs, out.module, "cg_pairs") != hipSuccess ||
        hipModuleGetFunction(&out.chain, out.module, "cg_chain") != hipSuccess)
        throw std::runtime_error("generated kernels not found in the module");
}

}  // namespace gbi
