// Sequence compiler: a PSD XML (MRiLab's MRiLabSeq format) and the VCtl
// scan parameters to the VSeq lines the scan engines take, without MATLAB.
//
// MRiLab turns a PSD XML into a MATLAB function (DoWriteXML2m.m:11-337),
// which DoPulseGen.m:18-117 calls once per TR and concatenates; each
// sequence element is a function of Macro/SeqElem. This file restates that
// pipeline in C++:
//  - XmlReader: elements, attributes and nesting of the PSD file;
//  - Expr: the attribute expressions (numbers, CVn, VCtl.x, VCtl.x(i),
//    VVar.x, VObj.Gyro, pi, Inf, + - * / ^, unary signs, parentheses, abs,
//    rem, mod, floor, ceil, round, sqrt, exp, max, min), evaluated in fp64
//    left to right as MATLAB does; "$k'a','b',..." picks option k;
//  - the elements rfSinc, rfRect, rfGaussian, rfFermi, GzSelective,
//    GzCartesian, GyCartesian, GxCartesian, ADCCartesian and ExtBit, with
//    StdTrap/StdRect/StdRamp/StdBit, DoB1Scaling and MATLAB's linspace;
//  - psd_tr: one TR of the generated PSD function (section gating by
//    TRStart/TREnd/Freq/Switch, rf TE anchoring, master-coil selection,
//    clipping to the section, |rf| < eps to 0, time offset, then the
//    stable sort of the merged time points, unique, and zeroed first and
//    last amplitudes);
//  - compile: DoPulseGen's slice x phase loop with dummy pulses and line
//    concatenation.
// Times are seconds, gradients T/m, rf amplitudes T, as in MRiLab.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "mrilab_amd.h"

namespace mrs {
// engine.hip: records msg as mrs_last_error() and returns st
mrs_status fail(mrs_status st, const std::string &msg);
}  // namespace mrs

struct mrs_seq {
    std::vector<double> uts, ts, rfAmp, rfPhase, rfFreq, rfCoil, Gz, Gy, Gx, ADC, Ext, flags;
    int32_t n_tr = 0;
    bool multi_transmit = false;  // VCtl.MultiTransmit == 'on'
    double master_tx = 1;         // VCtl.MasterTxCoil
};

namespace mrs {
namespace seqgen {

using Vec = std::vector<double>;
constexpr double kPi = 3.141592653589793;  // MATLAB's pi
constexpr double kEps = 2.220446049250313e-16;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- XML
struct Node {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Node> kids;
    const std::string *attr(const std::string &n) const {
        for (auto &a : attrs)
            if (a.first == n) return &a.second;
        return nullptr;
    }
};

class XmlReader {
   public:
    explicit XmlReader(const std::string &s) : s_(s) {}
    Node document() {
        skip_misc();
        if (i_ >= s_.size() || s_[i_] != '<') throw Error("PSD XML: no root element");
        Node root = element();
        skip_misc();
        if (i_ != s_.size()) throw Error("PSD XML: text after the root element");
        return root;
    }

   private:
    const std::string &s_;
    size_t i_ = 0;
    bool starts(const char *p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
    void skip_ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
    }
    void skip_until(const char *end) {
        const size_t k = s_.find(end, i_);
        if (k == std::string::npos) throw Error(std::string("PSD XML: missing ") + end);
        i_ = k + std::strlen(end);
    }
    void skip_misc() {  // whitespace, <?...?>, <!--...-->, <!DOCTYPE ...>
        for (;;) {
            skip_ws();
            if (starts("<?"))
                skip_until("?>");
            else if (starts("<!--"))
                skip_until("-->");
            else if (starts("<!"))
                skip_until(">");
            else
                return;
        }
    }
    std::string name() {
        const size_t b = i_;
        while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == ':' ||
                                  s_[i_] == '-' || s_[i_] == '.'))
            i_++;
        if (b == i_) throw Error("PSD XML: expected a name at offset " + std::to_string(b));
        return s_.substr(b, i_ - b);
    }
    static std::string unescape(const std::string &v) {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] != '&') {
                o += v[k];
                continue;
            }
            const size_t e = v.find(';', k);
            if (e == std::string::npos) throw Error("PSD XML: bad entity");
            const std::string ent = v.substr(k + 1, e - k - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else throw Error("PSD XML: unknown entity &" + ent + ";");
            k = e;
        }
        return o;
    }
    Node element() {
        i_++;  // '<'
        Node n;
        n.name = name();
        for (;;) {
            skip_ws();
            if (i_ >= s_.size()) throw Error("PSD XML: unterminated tag <" + n.name);
            if (starts("/>")) {
                i_ += 2;
                return n;
            }
            if (s_[i_] == '>') {
                i_++;
                break;
            }
            std::string an = name();
            skip_ws();
            if (i_ >= s_.size() || s_[i_] != '=') throw Error("PSD XML: attribute " + an + " has no value");
            i_++;
            skip_ws();
            const char q = i_ < s_.size() ? s_[i_] : 0;
            if (q != '"' && q != '\'') throw Error("PSD XML: unquoted attribute " + an);
            const size_t e = s_.find(q, i_ + 1);
            if (e == std::string::npos) throw Error("PSD XML: unterminated attribute " + an);
            n.attrs.emplace_back(an, unescape(s_.substr(i_ + 1, e - i_ - 1)));
            i_ = e + 1;
        }
        for (;;) {  // content: child elements; text is ignored
            skip_misc();
            if (i_ >= s_.size()) throw Error("PSD XML: missing </" + n.name + ">");
            if (starts("</")) {
                i_ += 2;
                if (name() != n.name) throw Error("PSD XML: mismatched </...> for <" + n.name + ">");
                skip_ws();
                if (i_ >= s_.size() || s_[i_] != '>') throw Error("PSD XML: bad closing tag");
                i_++;
                return n;
            }
            if (s_[i_] == '<') {
                n.kids.push_back(element());
            } else {
                while (i_ < s_.size() && s_[i_] != '<') i_++;
            }
        }
    }
};

// ---------------------------------------------------------------- values
struct Ctx {
    std::map<std::string, Vec> num;          // VCtl numbers (arrays allowed)
    std::map<std::string, std::string> txt;  // VCtl text fields
    std::map<std::string, double> cv;        // CVs of the PSD
    double tr_count = 0, slice_count = 0, phase_count = 0, gyro = 0;

    double vctl(const std::string &f, int idx = 1) const {
        auto it = num.find(f);
        if (it == num.end()) throw Error("VCtl." + f + " is not set");
        if (idx < 1 || idx > (int)it->second.size())
            throw Error("VCtl." + f + "(" + std::to_string(idx) + ") is out of range");
        return it->second[idx - 1];
    }
    std::string vtext(const std::string &f, const std::string &dflt) const {
        auto it = txt.find(f);
        return it == txt.end() ? dflt : it->second;
    }
    bool has(const std::string &f) const { return num.count(f) > 0; }
};

// Recursive-descent evaluator of a MATLAB scalar expression, left to right
// within a precedence level (MATLAB: ^ binds tighter than unary minus).
class Expr {
   public:
    Expr(const std::string &s, const Ctx &c) : s_(s), c_(c) {}
    double run() {
        const double v = sum();
        ws();
        if (i_ != s_.size()) throw Error("cannot evaluate '" + s_ + "'");
        return v;
    }

   private:
    const std::string &s_;
    const Ctx &c_;
    size_t i_ = 0;
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
    }
    bool eat(char ch) {
        ws();
        if (i_ < s_.size() && s_[i_] == ch) {
            i_++;
            return true;
        }
        return false;
    }
    double sum() {
        double v = prod();
        for (;;) {
            if (eat('+'))
                v = v + prod();
            else if (eat('-'))
                v = v - prod();
            else
                return v;
        }
    }
    double prod() {
        double v = unary();
        for (;;) {
            if (eat('*'))
                v = v * unary();
            else if (eat('/'))
                v = v / unary();
            else
                return v;
        }
    }
    double unary() {
        if (eat('-')) return -unary();
        if (eat('+')) return unary();
        return power();
    }
    double power() {
        double v = primary();
        while (eat('^')) {
            double e;
            if (eat('-'))
                e = -primary();
            else {
                eat('+');
                e = primary();
            }
            v = std::pow(v, e);
        }
        return v;
    }
    std::string ident() {
        ws();
        const size_t b = i_;
        while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_')) i_++;
        return s_.substr(b, i_ - b);
    }
    std::vector<double> args() {
        std::vector<double> a;
        if (eat(')')) return a;
        for (;;) {
            a.push_back(sum());
            if (eat(')')) return a;
            if (!eat(',')) throw Error("expected ',' or ')' in '" + s_ + "'");
        }
    }
    double primary() {
        ws();
        if (i_ >= s_.size()) throw Error("unexpected end of '" + s_ + "'");
        const char ch = s_[i_];
        if (ch == '(') {
            i_++;
            const double v = sum();
            if (!eat(')')) throw Error("missing ')' in '" + s_ + "'");
            return v;
        }
        if (std::isdigit((unsigned char)ch) || ch == '.') {
            const char *p = s_.c_str() + i_;
            char *end = nullptr;
            const double v = std::strtod(p, &end);
            if (end == p) throw Error("bad number in '" + s_ + "'");
            i_ += (size_t)(end - p);
            return v;
        }
        std::string id = ident();
        if (id.empty()) throw Error("cannot evaluate '" + s_ + "'");
        if (id == "VCtl" || id == "VVar" || id == "VObj") {
            if (!eat('.')) throw Error("expected a field after " + id);
            const std::string f = ident();
            if (id == "VVar") {
                if (f == "TRCount") return c_.tr_count;
                if (f == "SliceCount") return c_.slice_count;
                if (f == "PhaseCount") return c_.phase_count;
                throw Error("VVar." + f + " is not available to a PSD");
            }
            if (id == "VObj") {
                if (f == "Gyro") return c_.gyro;
                throw Error("VObj." + f + " is not available to a PSD");
            }
            if (eat('(')) {
                const auto a = args();
                if (a.size() != 1 || a[0] != std::floor(a[0])) throw Error("bad index of VCtl." + f);
                return c_.vctl(f, (int)a[0]);
            }
            return c_.vctl(f);
        }
        if (id == "pi") return kPi;
        if (id == "Inf" || id == "inf") return INFINITY;
        if (id == "NaN" || id == "nan") return NAN;
        if (id == "eps") return kEps;
        auto cv = c_.cv.find(id);
        if (cv != c_.cv.end()) return cv->second;
        if (eat('(')) {
            const auto a = args();
            auto one = [&](void) {
                if (a.size() != 1) throw Error(id + " takes one argument");
                return a[0];
            };
            auto two = [&](void) {
                if (a.size() != 2) throw Error(id + " takes two arguments");
            };
            if (id == "abs") return std::fabs(one());
            if (id == "floor") return std::floor(one());
            if (id == "ceil") return std::ceil(one());
            if (id == "round") return std::round(one());  // half away from zero, as MATLAB
            if (id == "sqrt") return std::sqrt(one());
            if (id == "exp") return std::exp(one());
            if (id == "rem") {
                two();
                return a[1] == 0 ? NAN : std::fmod(a[0], a[1]);
            }
            if (id == "mod") {
                two();
                if (a[1] == 0) return a[0];
                const double r = std::fmod(a[0], a[1]);
                return (r != 0 && ((r < 0) != (a[1] < 0))) ? r + a[1] : r;
            }
            if (id == "max" || id == "min") {
                two();
                return id == "max" ? std::max(a[0], a[1]) : std::min(a[0], a[1]);
            }
            throw Error("unknown function " + id + " in '" + s_ + "'");
        }
        throw Error("unknown name " + id + " in '" + s_ + "'");
    }
};

// An attribute as DoWriteXML2m.m:219-232 turns it into MATLAB: "$k'a','b'"
// selects option k, Notes is text, anything else an expression.
struct Param {
    bool text = false, empty = false;
    double v = 0;
    std::string s;
};

static std::vector<std::string> option_list(const std::string &v) {
    std::vector<std::string> out;
    size_t k = 0;
    while (k < v.size()) {
        while (k < v.size() && (std::isspace((unsigned char)v[k]) || v[k] == ',')) k++;
        if (k >= v.size()) break;
        if (v[k] != '\'') throw Error("bad option list '" + v + "'");
        const size_t e = v.find('\'', k + 1);
        if (e == std::string::npos) throw Error("bad option list '" + v + "'");
        out.push_back(v.substr(k + 1, e - k - 1));
        k = e + 1;
    }
    return out;
}

static Param param_of(const std::string &name, const std::string &v, const Ctx &c) {
    Param p;
    if (v.empty()) {
        p.empty = true;
        return p;
    }
    if (v[0] == '$') {
        if (v.size() < 2 || !std::isdigit((unsigned char)v[1])) throw Error("bad option attribute " + name);
        const auto opts = option_list(v.substr(2));
        const int k = v[1] - '0';
        if (k < 1 || k > (int)opts.size()) throw Error("option index out of range in " + name);
        p.text = true;
        p.s = opts[k - 1];
        return p;
    }
    if (name == "Notes") {
        p.text = true;
        p.s = v;
        return p;
    }
    p.v = Expr(v, c).run();
    return p;
}

struct Params {
    std::map<std::string, Param> m;
    std::string elem;
    double num(const std::string &n) const {
        auto it = m.find(n);
        if (it == m.end() || it->second.empty) throw Error(elem + "." + n + " is not set");
        if (it->second.text) throw Error(elem + "." + n + " must be numeric");
        return it->second.v;
    }
    std::string text(const std::string &n) const {
        auto it = m.find(n);
        if (it == m.end() || it->second.empty) throw Error(elem + "." + n + " is not set");
        if (!it->second.text) throw Error(elem + "." + n + " must be an option");
        return it->second.s;
    }
};

// ---------------------------------------------------------------- helpers
// Largest sample count one element may produce: a bad dt, Duplicates or
// spiral step (zero, NaN, huge) becomes an error instead of an allocation
// of unbounded size.
constexpr double kMaxPoints = 1 << 28;

static int64_t checked_count(double nd, const char *what) {
    if (!(nd < kMaxPoints)) throw Error(std::string(what) + ": sample count not finite or above 2^28");
    return (int64_t)std::floor(nd);
}

static int echo_train(double etl) { return (int)checked_count(etl, "echo train length"); }

// MATLAB linspace: d1 + (k*(d2 - d1))/(n - 1), with the ends set exactly.
static Vec linspace(double d1, double d2, double nd) {
    const int64_t n = checked_count(nd, "linspace");
    if (n < 1) return Vec();
    if (n == 1) return Vec{d2};
    Vec y((size_t)n);
    const double n1 = (double)(n - 1), d = d2 - d1;
    for (int64_t k = 0; k < n; k++) y[(size_t)k] = d1 + ((double)k * d) / n1;
    y[0] = d1;
    y[(size_t)n - 1] = d2;
    return y;
}

// [t, m] = unique(t) (sorted, first occurrence) applied to (amp, t).
static void unique_pair(Vec &amp, Vec &t) {
    std::vector<size_t> ix(t.size());
    std::iota(ix.begin(), ix.end(), 0);
    std::stable_sort(ix.begin(), ix.end(), [&](size_t a, size_t b) { return t[a] < t[b]; });
    Vec a2, t2;
    for (size_t k = 0; k < ix.size(); k++) {
        if (k > 0 && t[ix[k]] == t2.back()) continue;
        t2.push_back(t[ix[k]]);
        a2.push_back(amp[ix[k]]);
    }
    amp.swap(a2);
    t.swap(t2);
}

static void append(Vec &a, const Vec &b) { a.insert(a.end(), b.begin(), b.end()); }

// StdTrap.m: ramp up, flat top, ramp down; then unique on time.
static void std_trap(double t0, double t1, double tu, double td, double amp, double su, double sf, double sd,
                     Vec &g, Vec &t) {
    g.clear();
    t.clear();
    append(t, linspace(t0, tu, su));
    append(g, linspace(0.0, amp, su));
    const Vec tf = linspace(tu, td, sf);
    append(t, tf);
    g.insert(g.end(), tf.size(), amp);
    append(t, linspace(td, t1, sd));
    append(g, linspace(amp, 0.0, sd));
    unique_pair(g, t);
}

// DoB1Scaling.m: gain giving flip angle FA with the decoder's min-hold rule.
static double b1_gain(const Vec &p, double dt, double fa, double gyro) {
    const size_t n = p.size();
    std::vector<char> flag(n, 0);
    for (size_t k = 0; k + 1 < n; k++) {
        const double d = std::fabs(p[k + 1]) - std::fabs(p[k]);
        if (d < 0) flag[k + 1] = 1;
        if (d >= 0) flag[k] = 1;
    }
    double s = 0;
    for (size_t k = 0; k < n; k++) s = s + (flag[k] ? p[k] : 0.0);
    return ((fa / 180) * kPi) / (gyro * s * dt);
}

// Duplicates/DupSpacing of every element: repeat the samples, shifting the
// copies by DupSpacing.
static void duplicate(std::vector<Vec *> vals, Vec &t, const Params &p) {
    const double dups = std::max(1.0, p.num("Duplicates")), sp = std::max(0.0, p.num("DupSpacing"));
    if (dups == 1 || sp == 0) return;
    const int64_t nd = checked_count(dups * (double)std::max<size_t>(t.size(), 1), "Duplicates") > 0
                           ? (int64_t)dups : 1;
    for (Vec *v : vals) {
        Vec o;
        for (int64_t d = 0; d < nd; d++) append(o, *v);
        v->swap(o);
    }
    Vec o;
    for (int64_t d = 0; d < nd; d++)
        for (double x : t) o.push_back(x + (double)d * sp);
    t.swap(o);
}

// ---------------------------------------------------------------- elements
struct RfOut {
    Vec amp, phase, freq, coil, t;
};

static void rf_finish(RfOut &r, const Params &p, const Ctx &c) {
    const size_t n = r.t.size();
    const double dt = p.num("dt");
    r.amp.front() = 0;
    r.amp.back() = 0;
    const double g = b1_gain(r.amp, dt, p.num("FA"), c.gyro);
    for (double &a : r.amp) a = g * a;
    r.phase.assign(n, p.num("rfPhase"));
    r.freq.assign(n, p.num("rfFreq"));
    r.coil.assign(n, p.num("CoilID"));
    r.phase.front() = r.phase.back() = 0;
    r.freq.front() = r.freq.back() = 0;
    duplicate({&r.amp, &r.freq, &r.phase, &r.coil}, r.t, p);
}

static bool adiabatic(const std::string &name) {
    return name == "rfHyperbolicSecant" || name == "rfTanhTan" || name == "rfBIR" || name == "rfBIREF";
}

// The adiabatic pulses (rfHyperbolicSecant.m, rfTanhTan.m, rfBIR.m,
// rfBIREF.m): amplitude MaxB1 (no flip-angle scaling), frequency sweep and
// phase over one or more segments of the pulse; ends zeroed, duplicated.
static RfOut rf_adiabatic(const std::string &name, const Params &p) {
    RfOut r;
    const double ts = p.num("tStart"), te = p.num("tEnd"), dt = p.num("dt"), b1 = p.num("MaxB1");
    const double T = te - ts;  // the pulse on [0, T], shifted by tStart at the end
    if (name == "rfHyperbolicSecant") {
        r.t = linspace(ts, te, std::ceil((te - ts) / dt) + 1);
    } else {
        r.t = linspace(0, T, std::ceil(T / dt) + 1);
    }
    const size_t n = r.t.size();
    if (n < 2) throw Error(name + ": fewer than 2 samples");
    r.amp.assign(n, 0.0);
    r.freq.assign(n, 0.0);
    r.phase.assign(n, 0.0);
    if (name == "rfHyperbolicSecant") {  // rfHyperbolicSecant.m:27-31
        const double tbp = p.num("TBP"), ad = p.num("Adiab"), beta = (tbp / (te - ts)) * kPi / ad;
        for (size_t k = 0; k < n; k++) {
            const double x = beta * (r.t[k] - (te - ts) / 2 - ts);
            r.amp[k] = b1 * (1 / std::cosh(x));
            r.freq[k] = -beta * ad * std::tanh(x) / (2 * kPi);
        }
        r.phase.assign(n, p.num("rfPhase"));
    } else if (name == "rfTanhTan") {  // rfTanhTan.m:25-40
        const double zeta = 10, kappa = std::atan(20.0), a = p.num("TBP") * kPi / T;
        for (size_t k = 0; k < n; k++) {
            const bool first = r.t[k] < T / 2;
            const double u = first ? r.t[k] : T - r.t[k];
            r.amp[k] = b1 * std::tanh((2 * zeta * u) / T);
            const double f = (std::tan(kappa * (1 - 2 * u / T)) / std::tan(kappa));
            r.freq[k] = (first ? a : -a) * f / (2 * kPi);
        }
        r.phase.assign(n, p.num("rfPhase"));
    } else if (name == "rfBIR") {  // rfBIR.m:28-79
        const std::string flag = p.text("BIRFlag");
        const double mf = p.num("MaxFreq");
        if (flag == "BIR-1") {
            const double z = kPi / T;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k];
                const bool first = t < T / 2;
                r.freq[k] = (first ? mf : -mf) * std::sin(z * t);
                r.phase[k] = first ? 0.0 : kPi / 2;
                r.amp[k] = b1 * std::cos(z * t);
            }
        } else if (flag == "BIR-2") {
            const double z = 2 * kPi / T;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k];
                r.phase[k] = t < T / 4 ? 0.0 : t < T / 2 ? kPi / 2 : kPi / 2 + kPi;
                r.amp[k] = std::fabs(b1 * std::cos(z * t));
                r.freq[k] = std::fabs(mf * std::sin(z * t));
            }
        } else if (flag == "BIR-4") {
            const double lam = p.num("Lambda"), beta = p.num("Beta"), den = std::tan(beta) * 2 * kPi;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k], q = 4 * t / T;
                if (t < T / 4) {
                    r.amp[k] = b1 * std::tanh(lam * (1 - q));
                    r.freq[k] = std::tan(4 * beta * t / T) / den;
                } else if (t < T / 2) {
                    r.amp[k] = b1 * std::tanh(lam * (q - 1));
                    r.freq[k] = std::tan(beta * (q - 2)) / den;
                } else if (t < 3 * T / 4) {
                    r.amp[k] = b1 * std::tanh(lam * (3 - q));
                    r.freq[k] = std::tan(beta * (q - 2)) / den;
                } else {
                    r.amp[k] = b1 * std::tanh(lam * (q - 3));
                    r.freq[k] = std::tan(beta * (q - 4)) / den;
                }
            }
        } else {
            throw Error("rfBIR: unknown BIRFlag " + flag);
        }
    } else {  // rfBIREF.m:27-72
        const std::string flag = p.text("BIREFFlag");
        const double mf = p.num("MaxFreq");
        if (flag == "BIREF-1") {
            const double z = kPi / T;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k];
                r.amp[k] = (t < T / 2 ? b1 : -b1) * std::sin(z * t);
                r.freq[k] = mf * std::fabs(std::cos(z * t));
            }
        } else if (flag == "BIREF-2a") {
            const double z = kPi / T;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k];
                r.freq[k] = (t < T / 2 ? mf : -mf) * std::sin(z * t);
                r.amp[k] = b1 * std::fabs(std::cos(z * t));
            }
        } else if (flag == "BIREF-2b") {
            const double z = 2 * kPi / T;
            for (size_t k = 0; k < n; k++) {
                const double t = r.t[k];
                if (t < T / 2) {
                    r.amp[k] = b1 * std::fabs(std::cos(z * t));
                    r.freq[k] = (t < T / 4 ? mf : -mf) * std::sin(z * t);
                } else {
                    r.amp[k] = -b1 * std::cos(z * t);
                    r.freq[k] = (t < 3 * T / 4 ? -mf : mf) * std::sin(z * t);
                }
            }
        } else {
            throw Error("rfBIREF: unknown BIREFFlag " + flag);
        }
    }
    if (name != "rfHyperbolicSecant")
        for (double &t : r.t) t = t + ts;
    r.coil.assign(n, p.num("CoilID"));
    r.amp.front() = r.amp.back() = 0;
    r.freq.front() = r.freq.back() = 0;
    r.phase.front() = r.phase.back() = 0;
    duplicate({&r.amp, &r.freq, &r.phase, &r.coil}, r.t, p);
    return r;
}

static RfOut rf_element(const std::string &name, const Params &p, const Ctx &c) {
    if (adiabatic(name)) return rf_adiabatic(name, p);
    RfOut r;
    const double ts = p.num("tStart"), te = p.num("tEnd"), dt = p.num("dt");
    r.t = linspace(ts, te, std::ceil((te - ts) / dt) + 1);
    const size_t n = r.t.size();
    if (n < 2) throw Error(name + ": fewer than 2 samples");
    r.amp.resize(n);
    if (name == "rfSinc") {  // rfSinc.m:24-42
        const std::string apod = p.text("Apod");
        const double a = apod == "Hamming" ? 0.46 : apod == "Hanning" ? 0.5 : 0.0;
        if (apod != "Non" && apod != "Hamming" && apod != "Hanning") throw Error("rfSinc: unknown Apod " + apod);
        const double tbp = p.num("TBP"), t0 = (te - ts) / tbp;
        double mx = -INFINITY;
        std::vector<char> nan(n, 0);
        for (size_t k = 0; k < n; k++) {
            const double t = r.t[k] - (te - ts) / 2 - ts;  // centred, then shifted back below
            const double v = ((t0 * std::sin(kPi * t / t0)) / (kPi * t)) * ((1 - a) + a * std::cos((kPi * t) / (tbp * t0)));
            r.amp[k] = v;
            r.t[k] = t + (te - ts) / 2 + ts;
            if (std::isnan(v))
                nan[k] = 1;
            else
                mx = std::max(mx, v);
        }
        for (size_t k = 0; k < n; k++)
            if (nan[k]) r.amp[k] = mx;
    } else if (name == "rfRect") {  // rfRect.m
        std::fill(r.amp.begin(), r.amp.end(), 1.0);
    } else if (name == "rfGaussian") {  // rfGaussian.m:19-23
        const double sigma = (te - ts) / 7.434;
        for (size_t k = 0; k < n; k++) {
            const double d = r.t[k] - (te - ts) / 2 - ts;
            r.amp[k] = std::exp(-(d * d) / (2 * (sigma * sigma)));
        }
    } else if (name == "rfFermi") {  // rfFermi.m:20-23
        const double pw = p.num("PW"), tw = ((te - ts) - 2 * pw) / 13.81;
        for (size_t k = 0; k < n; k++)
            r.amp[k] = 1 / (1 + std::exp((std::fabs(r.t[k] - (te - ts) / 2 - ts) - pw) / tw));
    } else {
        throw Error("rf element " + name + " is not supported");
    }
    rf_finish(r, p, c);
    return r;
}

struct GOut {
    Vec amp, t;
};

// GzCartesian.m / GyCartesian.m: phase-encoding trapezoid and optional rephaser.
static GOut g_cartesian_pe(const Params &p, const Ctx &c, bool z) {
    const double t1s = p.num("t1Start"), t1e = p.num("t1End"), t2s = p.num("t2Start"), t2e = p.num("t2End");
    const double tr = p.num("tRamp");
    const double s1 = p.num(z ? "Gz1Sign" : "Gy1Sign"), s2 = p.num(z ? "Gz2Sign" : "Gy2Sign");
    const double fov = c.vctl(z ? "FOVSlice" : "FOVPhase");
    const double step = z ? (c.slice_count - std::floor(c.vctl("SliceNum") / 2) - 1)
                          : (c.phase_count - c.vctl("ResPhase") / 2 - 1);
    const double gbar = c.gyro / (2 * kPi);
    GOut o;
    Vec g2, t2;
    double dG = 1 / (gbar * (t1e - t1s) * fov);
    std_trap(t1s - tr, t1e + tr, t1s, t1e, step * dG * s1, 2, 2, 2, o.amp, o.t);
    dG = 1 / (gbar * (t2e - t2s) * fov);
    std_trap(t2s - tr, t2e + tr, t2s, t2e, step * dG * s2, 2, 2, 2, g2, t2);
    if (s2 != 0) {
        append(o.amp, g2);
        append(o.t, t2);
    }
    unique_pair(o.amp, o.t);
    duplicate({&o.amp}, o.t, p);
    return o;
}

// GxCartesian.m: prephaser, readout centred at t2Middle + TEAnchorTime, rewinder.
static GOut gx_cartesian(const Params &p, const Ctx &c) {
    const double t1s = p.num("t1Start"), t2m = p.num("t2Middle"), t3s = p.num("t3Start"), tr = p.num("tRamp");
    const double s1 = p.num("Gx1Sign"), s2 = p.num("Gx2Sign"), s3 = p.num("Gx3Sign");
    const double gbar = c.gyro / (2 * kPi);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq"));
    const double ta = c.vctl("TEAnchorTime");
    Vec g1, tt1, g2, tt2, g3, tt3;
    std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, gx * s1, 2, 2, 2, g1, tt1);
    std_trap(t2m + ta - th - tr, t2m + ta + th + tr, t2m + ta - th, t2m + ta + th, gx * s2, 2, 2, 2, g2, tt2);
    std_trap(t3s - tr, t3s + th + tr, t3s, t3s + th, gx * s3, 2, 2, 2, g3, tt3);
    GOut o;
    if (s1 != 0) {
        o.amp = g1;
        o.t = tt1;
    }
    append(o.amp, g2);
    append(o.t, tt2);
    if (s3 != 0) {
        append(o.amp, g3);
        append(o.t, tt3);
    }
    unique_pair(o.amp, o.t);
    duplicate({&o.amp}, o.t, p);
    return o;
}

// GzSelective.m: slice-select trapezoid with optional pre- and rephasers.
static GOut gz_selective(const Params &p) {
    const double t2s = p.num("t2Start"), t2e = p.num("t2End"), tr = p.num("tRamp"), ga = p.num("GzAmp");
    const double s1 = p.num("Gz1Sign"), s2 = p.num("Gz2Sign"), s3 = p.num("Gz3Sign");
    Vec g1, tt1, g2, tt2, g3, tt3;
    const double half = (t2e - t2s) / 2;
    std_trap(0, half + 2 * tr, tr, half + tr, s1 * ga, 2, 2, 2, g1, tt1);
    std_trap(t2s - tr, t2e + tr, t2s, t2e, s2 * ga, 2, 2, 2, g2, tt2);
    std_trap(0, half + 2 * tr, tr, half + tr, s3 * ga, 2, 2, 2, g3, tt3);
    GOut o;
    if (s1 != 0) {
        o.amp = g1;
        const double last = tt1.back();
        for (double t : tt1) o.t.push_back(t2s - tr - last + t);
    }
    append(o.amp, g2);
    append(o.t, tt2);
    if (s3 != 0) {
        append(o.amp, g3);
        for (double t : tt3) o.t.push_back(t2e + tr + t);
    }
    unique_pair(o.amp, o.t);
    duplicate({&o.amp}, o.t, p);
    return o;
}

// ADCCartesian.m: ResFreq samples over the readout flat top.
static GOut adc_cartesian(const Params &p, const Ctx &c) {
    const double tm = p.num("tMiddle");
    const double gbar = c.gyro / (2 * kPi);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq"));
    const double ta = c.vctl("TEAnchorTime"), mu = c.vctl("MinUpdRate");
    GOut o;
    std_trap(tm + ta - th - mu, tm + ta + th + mu, tm + ta - th, tm + ta + th, 1, 2, c.vctl("ResFreq"), 2, o.amp, o.t);
    unique_pair(o.amp, o.t);
    duplicate({&o.amp}, o.t, p);
    return o;
}

// ------------------------------------------------- EPI, FSE, radial, spiral
static double sgn(double x) { return x > 0 ? 1.0 : x < 0 ? -1.0 : 0.0; }

static bool echo_shift(const Ctx &c) { return c.vtext("EPI_EchoShifting", "off") == "on"; }

// GxAreaTrapezoid2.m / GyAreaTrapezoid2.m: area (1/m) at the largest
// gradient MaxGrad, ramps at MaxSlewRate, at least MinUpdRate long.
static GOut area_trap2(double ts, double area, const Ctx &c) {
    const double a = std::fabs(area) / (c.gyro / (2 * kPi)), mu = c.vctl("MinUpdRate");
    double amp = c.vctl("MaxGrad"), te;
    if (a / amp < mu) {
        te = mu + ts;
        amp = a / mu;
    } else {
        te = a / amp + ts;
    }
    const double tr = std::max(mu, amp / c.vctl("MaxSlewRate"));
    GOut o;
    std_trap(ts - tr, te + tr, ts, te, amp * sgn(area), 2, 2, 2, o.amp, o.t);
    return o;
}

// GyAreaTrapezoid.m: area (1/m) over a fixed flat top [ts, te].
static GOut area_trap(double ts, double te, double area, const Ctx &c) {
    const double a = std::fabs(area) / (c.gyro / (2 * kPi));
    const double amp = a / (te - ts), tr = std::max(c.vctl("MinUpdRate"), amp / c.vctl("MaxSlewRate"));
    GOut o;
    std_trap(ts - tr, te + tr, ts, te, amp * sgn(area), 2, 2, 2, o.amp, o.t);
    return o;
}

// ADCBlock.m: sSample samples over [ts, te] (one flagged point if 1).
static GOut adc_block(double ts, double te, double n, const Ctx &c) {
    GOut o;
    if (n == 1) {
        o.t = linspace(ts, te, 3);
        o.amp = {0.0, 1.0, 0.0};
    } else {
        const double mu = c.vctl("MinUpdRate");
        std_trap(ts - mu, te + mu, ts, te, 1, 2, n, 2, o.amp, o.t);
    }
    return o;
}

static void cat(GOut &o, const GOut &g, double scale = 1.0) {
    for (double a : g.amp) o.amp.push_back(a * scale);
    append(o.t, g.t);
}

// GxEPI.m: area prephaser, then EPI_ETL readout lobes of alternating sign.
static GOut gx_epi(const Params &p, const Ctx &c) {
    const double t1s = p.num("t1Start"), t2m = p.num("t2Middle"), s1 = p.num("Gx1Sign"), s2 = p.num("Gx2Sign");
    const double gbar = c.gyro / (2 * kPi);
    const GOut pre = area_trap2(t1s, (1 / c.vctl("RFreq")) / 2, c);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq"));
    const double esp = c.vctl("EPI_ESP"), etl = c.vctl("EPI_ETL");
    double off = (t2m + c.vctl("TEAnchorTime")) - (std::floor(etl / 2) + 0.5) * esp;
    if (echo_shift(c)) off = off + (c.phase_count - 1) * (esp / c.vctl("EPI_ShotNum"));
    const double tr = esp / 2 - th;
    GOut o;
    cat(o, pre, s1);
    for (int i = 1; i <= echo_train(etl); i++) {
        const double ts = off + (i - 1) * esp + tr, te = ts + 2 * th;
        GOut g;
        std_trap(ts - tr, te + tr, ts, te, gx, 2, 2, 2, g.amp, g.t);
        cat(o, g, ((i - 1) % 2 ? -1.0 : 1.0) * s2);
    }
    unique_pair(o.amp, o.t);
    return o;
}

// GyEPI.m: area prephaser to the shot's first line, then EPI_ETL - 1 blips.
static GOut gy_epi(const Params &p, const Ctx &c) {
    const double t1s = p.num("t1Start"), t2m = p.num("t2Middle"), s1 = p.num("Gy1Sign"), s2 = p.num("Gy2Sign");
    const double kmax = (1 / c.vctl("RPhase")) / 2, etl = c.vctl("EPI_ETL"), esp = c.vctl("EPI_ESP");
    const double dk = (2 * kmax) / etl;
    GOut o;
    cat(o, area_trap2(t1s, kmax - (c.phase_count - 1) * (dk / c.vctl("EPI_ShotNum")), c), s1);
    double off = (t2m + c.vctl("TEAnchorTime")) - (std::floor(etl / 2) - 0.5) * esp;
    if (echo_shift(c)) off = off + (c.phase_count - 1) * (esp / c.vctl("EPI_ShotNum"));
    for (int i = 1; i <= echo_train(etl) - 1; i++) cat(o, area_trap2(off + (i - 1) * esp, dk, c), s2);
    unique_pair(o.amp, o.t);
    return o;
}

// ADCEPI.m: ResFreq samples on the flat top of every readout lobe.
static GOut adc_epi(const Params &p, const Ctx &c) {
    const double tm = p.num("tMiddle"), gbar = c.gyro / (2 * kPi);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq"));
    const double esp = c.vctl("EPI_ESP"), etl = c.vctl("EPI_ETL");
    double off = (tm + c.vctl("TEAnchorTime")) - (std::floor(etl / 2) + 0.5) * esp;
    if (echo_shift(c)) off = off + (c.phase_count - 1) * (esp / c.vctl("EPI_ShotNum"));
    const double tr = esp / 2 - th;
    GOut o;
    for (int i = 1; i <= echo_train(etl); i++) {
        const double ts = off + (i - 1) * esp + tr;
        cat(o, adc_block(ts, ts + 2 * th, c.vctl("ResFreq"), c));
    }
    unique_pair(o.amp, o.t);
    return o;
}

// GxFSE.m: prephaser, then FSE_ETL readouts FSE_ESP apart, ramps at MaxSlewRate.
static GOut gx_fse(const Params &p, const Ctx &c) {
    const double t1s = p.num("t1Start"), t2m = p.num("t2Middle"), s1 = p.num("Gx1Sign"), s2 = p.num("Gx2Sign");
    const double gbar = c.gyro / (2 * kPi);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq")), tr = gx / c.vctl("MaxSlewRate");
    GOut o;
    std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, gx * s1, 2, 2, 2, o.amp, o.t);
    const double esp = c.vctl("FSE_ESP"), etl = c.vctl("FSE_ETL");
    const double off = (t2m + c.vctl("TEAnchorTime")) - std::floor(etl / 2) * esp;
    for (int i = 1; i <= echo_train(etl); i++) {
        const double m = off + (i - 1) * esp;
        GOut g;
        std_trap(m - th - tr, m + th + tr, m - th, m + th, gx * s2, 2, 2, 2, g.amp, g.t);
        cat(o, g);
    }
    unique_pair(o.amp, o.t);
    return o;
}

// GyFSE.m: per echo a phase encode before it and a rewinder after it.
static GOut gy_fse(const Params &p, const Ctx &c) {
    const double tm = p.num("tMiddle"), g1 = p.num("tGy1"), g2 = p.num("tGy2"), toff = p.num("tOffset");
    const double s1 = p.num("Gy1Sign"), s2 = p.num("Gy2Sign");
    const double kmax = (1 / c.vctl("RPhase")) / 2, etl = c.vctl("FSE_ETL"), esp = c.vctl("FSE_ESP");
    const double dk = (2 * kmax) / etl;
    const double off = (tm + c.vctl("TEAnchorTime")) - std::floor(etl / 2) * esp;
    GOut o;
    for (int i = 1; i <= echo_train(etl); i++) {
        const double k = -(kmax - (c.phase_count - 1) * (dk / c.vctl("FSE_ShotNum"))) + (i - 1) * dk;
        double ts = off + (i - 1) * esp - toff;
        cat(o, area_trap(ts, ts + g1, k * s1, c));
        ts = off + (i - 1) * esp + toff - g2;
        cat(o, area_trap(ts, ts + g2, k * s2, c));
    }
    unique_pair(o.amp, o.t);
    return o;
}

// ADCFSE.m: ResFreq samples per echo.
static GOut adc_fse(const Params &p, const Ctx &c) {
    const double tm = p.num("tMiddle"), gbar = c.gyro / (2 * kPi);
    const double gx = (1 / c.vctl("FOVFreq")) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * c.vctl("RFreq")), tr = c.vctl("MinUpdRate");
    const double esp = c.vctl("FSE_ESP"), etl = c.vctl("FSE_ETL");
    const double off = (tm + c.vctl("TEAnchorTime")) - std::floor(etl / 2) * esp;
    GOut o;
    for (int i = 1; i <= echo_train(etl); i++) {
        const double m = off + (i - 1) * esp;
        GOut g;
        std_trap(m - th - tr, m + th + tr, m - th, m + th, 1, 2, c.vctl("ResFreq"), 2, g.amp, g.t);
        cat(o, g);
    }
    unique_pair(o.amp, o.t);
    return o;
}

// GxRadial.m / GyRadial.m: the Cartesian readout rotated to spoke
// PhaseCount - 1 (cos for x, sin for y).
static GOut g_radial(const Params &p, const Ctx &c, bool y) {
    const double t1s = p.num("t1Start"), t2m = p.num("t2Middle"), t3s = p.num("t3Start"), tr = p.num("tRamp");
    const double s1 = p.num(y ? "Gy1Sign" : "Gx1Sign"), s2 = p.num(y ? "Gy2Sign" : "Gx2Sign"),
                 s3 = p.num(y ? "Gy3Sign" : "Gx3Sign");
    const double fov = c.vctl("FOVFreq"), res = c.vctl("ResFreq"), gbar = c.gyro / (2 * kPi);
    double inc;
    const std::string pat = c.vtext("R_AngPattern", "Linear");
    if (pat == "Linear") {
        const std::string range = c.vtext("R_AngRange", "");
        if (range.empty()) throw Error("VCtl.R_AngRange is not set");
        inc = Expr(range, c).run() / c.vctl("R_SpokeNum");
    } else if (pat == "Golden") {
        inc = 111.246 * (kPi / 180);
    } else {
        throw Error("VCtl.R_AngPattern must be Linear or Golden");
    }
    const double gx = (1 / fov) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * (fov / res));
    const double ang = inc * (c.phase_count - 1);
    const double amp = gx * (y ? std::sin(ang) : std::cos(ang)), ta = c.vctl("TEAnchorTime");
    Vec g1, tt1, g2, tt2, g3, tt3;
    std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, amp * s1, 2, 2, 2, g1, tt1);
    std_trap(t2m + ta - th - tr, t2m + ta + th + tr, t2m + ta - th, t2m + ta + th, amp * s2, 2, 2, 2, g2, tt2);
    std_trap(t3s - tr, t3s + th + tr, t3s, t3s + th, amp * s3, 2, 2, 2, g3, tt3);
    GOut o;
    if (s1 != 0) {
        o.amp = g1;
        o.t = tt1;
    }
    append(o.amp, g2);
    append(o.t, tt2);
    if (s3 != 0) {
        append(o.amp, g3);
        append(o.t, tt3);
    }
    unique_pair(o.amp, o.t);
    return o;
}

// ADCRadial.m: R_SampPerSpoke samples on the spoke's flat top.
static GOut adc_radial(const Params &p, const Ctx &c) {
    const double tm = p.num("tMiddle"), fov = c.vctl("FOVFreq"), res = c.vctl("ResFreq");
    const double gbar = c.gyro / (2 * kPi), gx = (1 / fov) / (gbar * (1 / c.vctl("BandWidth")));
    const double th = 1 / (2 * gbar * gx * (fov / res)), ta = c.vctl("TEAnchorTime"), mu = c.vctl("MinUpdRate");
    GOut o;
    std_trap(tm + ta - th - mu, tm + ta + th + mu, tm + ta - th, tm + ta + th, 1, 2, c.vctl("R_SampPerSpoke"), 2,
             o.amp, o.t);
    return o;
}

// a:d:b, with a relative tolerance on the end point (MATLAB's colon).
static Vec colon(double a, double d, double b) {
    const double q = (b - a) / d;
    if (!(q >= 0) || !std::isfinite(q)) return q == 0 ? Vec{a} : Vec();
    const int64_t n = checked_count(q + 1e-10, "spiral time grid") + 1;
    Vec y((size_t)n);
    for (int64_t k = 0; k < n; k++) y[(size_t)k] = a + (double)k * d;
    return y;
}

struct SpiralShape {
    Vec t, theta;
    double tacq;
};

// The variable-density spiral of GxSpiral.m / ADCSpiral.m: angle Theta(t)
// under the slew-rate limited start and the gradient-limited tail.
static SpiralShape spiral_shape(const Ctx &c, double dt, bool need_samples) {
    const double fov = c.vctl("FOVFreq"), res = c.vctl("ResFreq"), g = c.gyro;
    const double lam = c.vctl("S_Lamda"), sr = c.vctl("S_SlewRate"), grad = c.vctl("S_Gradient");
    const double kmax = res / (2 * fov), thmax = kmax / lam;
    const double beta = (g / (2 * kPi)) * (sr / lam);
    const double a2 = std::pow((9 * beta / 4), 1.0 / 3);
    const double cl = sr / c.vctl("S_SlewRate0");
    const double ts = std::pow(((3 * g * grad) / (4 * kPi * lam * (a2 * a2))), 3);
    auto th1 = [&](double t) { return (0.5 * beta * (t * t)) / (cl + (beta / (2 * a2)) * std::pow(t, 4.0 / 3)); };
    const double thts = th1(ts);
    SpiralShape s;
    if (thts < thmax) {
        Vec t1 = need_samples ? colon(0, dt, ts) : Vec{0.0, ts};
        const double t1e = need_samples ? t1.back() : ts;
        const double th1e = th1(t1e);
        s.tacq = t1e + ((kPi * lam) / (g * grad)) * ((thmax * thmax) - (th1e * th1e));
        if (need_samples) {
            for (size_t k = 0; k + 1 < t1.size(); k++) {
                s.t.push_back(t1[k]);
                s.theta.push_back(th1(t1[k]));
            }
            for (double t : colon(t1e, dt, s.tacq)) {
                s.t.push_back(t);
                s.theta.push_back(std::sqrt((th1e * th1e) + (g / (kPi * lam)) * grad * (t - t1e)));
            }
        }
    } else {
        s.tacq = ((2 * kPi * fov) / (3 * c.vctl("S_ShotNum"))) *
                 std::sqrt(kPi / (g * sr * std::pow(c.vctl("RFreq"), 3)));
        if (need_samples)
            for (double t : colon(0, dt, s.tacq)) {
                s.t.push_back(t);
                s.theta.push_back(th1(t));
            }
    }
    return s;
}

// GxSpiral.m / GySpiral.m: the spiral gradient of interleave PhaseCount.
static GOut g_spiral(const Params &p, const Ctx &c, bool y) {
    const double ts = p.num("tStart"), dt = p.num("dt");
    const SpiralShape s = spiral_shape(c, dt, true);
    const size_t n = s.t.size();
    if (n < 2) throw Error("spiral: fewer than 2 samples");
    const double rot = (c.phase_count - 1) * (2 * kPi / c.vctl("S_ShotNum"));
    const double lam = c.vctl("S_Lamda"), g = c.gyro;
    GOut o;
    for (size_t k = 0; k < n; k++) {
        const double dth = k == 0 ? 0.0 : (s.theta[k] - s.theta[k - 1]) / (s.t[k] - s.t[k - 1]);
        const double a = s.theta[k] + rot;
        const double f = y ? (std::sin(a) + s.theta[k] * std::cos(a)) : (std::cos(a) - s.theta[k] * std::sin(a));
        o.amp.push_back(((2 * kPi) / g) * lam * dth * f);
        o.t.push_back(ts + c.vctl("TEAnchorTime") + s.t[k]);
    }
    o.t.push_back(o.t.back() + std::max(c.vctl("MinUpdRate"), std::fabs(o.amp.back() / c.vctl("S_SlewRate"))));
    o.amp.push_back(0.0);
    unique_pair(o.amp, o.t);
    return o;
}

// ADCSpiral.m: floor(Tacq*BandWidth) samples over the spiral readout.
static GOut adc_spiral(const Params &p, const Ctx &c) {
    const double ts = p.num("tStart"), ta = c.vctl("TEAnchorTime"), mu = c.vctl("MinUpdRate");
    const double t = spiral_shape(c, 0, false).tacq;
    GOut o;
    std_trap(ts + ta - mu, ts + ta + t + mu, ts + ta, ts + ta + t, 1, 2, std::floor(t * c.vctl("BandWidth")), 2,
             o.amp, o.t);
    return o;
}

// {Gx,Gy,Gz}Trapezoid.m, {Gx,Gy,Gz}AreaTrapezoid.m and
// {Gx,Gy,Gz}AreaTrapezoid2.m as elements of their line (ax = "Gx", "Gy",
// "Gz"): a trapezoid of given amplitude, of given area over [tStart, tEnd],
// or of given area at MaxGrad; duplicated. False if name is none of them.
static bool g_trapezoid(const std::string &name, const std::string &ax, const Params &p, const Ctx &c, GOut &o) {
    if (name == ax + "Trapezoid") {
        const double ts = p.num("tStart"), te = p.num("tEnd"), tr = p.num("tRamp");
        const double sr = std::max(2.0, p.num("sRamp"));
        std_trap(ts - tr, te + tr, ts, te, p.num(ax + "Amp"), sr, 2, sr, o.amp, o.t);
    } else if (name == ax + "AreaTrapezoid") {
        o = area_trap(p.num("tStart"), p.num("tEnd"), p.num("Area"), c);
    } else if (name == ax + "AreaTrapezoid2") {
        o = area_trap2(p.num("tStart"), p.num("Area"), c);
    } else {
        return false;
    }
    duplicate({&o.amp}, o.t, p);
    return true;
}

// GzSelective2.m: a slice-select plateau [t2Start, t2End] with optional
// lobes of their own amplitude and duration before and after it, the lobe
// ramps running into the plateau's amplitude.
static GOut gz_selective2(const Params &p) {
    const double t2s = p.num("t2Start"), t2e = p.num("t2End"), tr = p.num("tRamp");
    const double tg1 = p.num("tGz1"), tg3 = p.num("tGz3");
    const double a1 = p.num("Gz1Amp"), a2 = p.num("Gz2Amp"), a3 = p.num("Gz3Amp");
    Vec g1, tt1, g3, tt3;
    std_trap(0, tg1 + 2 * tr, tr, tg1 + tr, a1, 2, 2, 2, g1, tt1);
    std_trap(0, tg3 + 2 * tr, tr, tg3 + tr, a3, 2, 2, 2, g3, tt3);
    const Vec t2 = linspace(t2s, t2e, 2);
    GOut o;
    if (a1 != 0 && tg1 != 0) {
        g1.back() = a2;
        o.amp = g1;
        for (double t : tt1) o.t.push_back(t2s - tt1.back() + t);
    } else {
        o.amp = {0.0};
        o.t = {t2[0] - tr};
    }
    o.amp.push_back(a2);
    o.amp.push_back(a2);
    append(o.t, t2);
    if (a3 != 0 && tg3 != 0) {
        g3.front() = a2;
        append(o.amp, g3);
        for (double t : tt3) o.t.push_back(t2e + t);
    } else {
        o.amp.push_back(0.0);
        o.t.push_back(o.t.back() + tr);
    }
    unique_pair(o.amp, o.t);
    duplicate({&o.amp}, o.t, p);
    return o;
}

// ExtBit.m / StdBit.m: one flagged point between two zeros.
static GOut ext_bit(const Params &p, const Ctx &c) {
    const double ts = p.num("tStart");
    GOut o;
    o.t = linspace(ts, ts + 2 * c.vctl("MinUpdRate"), 3);
    o.amp.assign(o.t.size(), p.num("Ext"));
    o.amp.front() = 0;
    o.amp.back() = 0;
    duplicate({&o.amp}, o.t, p);
    return o;
}

// ---------------------------------------------------------------- one TR
struct Lines {
    Vec rfAmp, rfPhase, rfFreq, rfCoil, rfT, Gz, GzT, Gy, GyT, Gx, GxT, ADC, ADCT, Ext, ExtT;
};

struct TrOut {
    Vec uts, ts, flags, rfAmp, rfPhase, rfFreq, rfCoil, Gz, Gy, Gx, ADC, Ext;
};

static Params params_of(const Node &e, const Ctx &c) {
    Params p;
    p.elem = e.name;
    for (auto &a : e.attrs) p.m[a.first] = param_of(a.first, a.second, c);
    return p;
}

static bool section_on(const Params &s, const Ctx &c) {
    const double start = s.num("TRStart"), end = s.num("TREnd"), freq = s.num("Freq");
    const double r = std::fmod(c.tr_count - start, freq);
    return !(c.tr_count < start || c.tr_count > end || r != 0 || s.text("Switch") == "off");
}

// Stable sort of a line by time, keeping the paired amplitudes.
static std::vector<size_t> order_of(const Vec &t) {
    std::vector<size_t> ix(t.size());
    std::iota(ix.begin(), ix.end(), 0);
    std::stable_sort(ix.begin(), ix.end(), [&](size_t a, size_t b) { return t[a] < t[b]; });
    return ix;
}

static Vec take(const Vec &v, const std::vector<size_t> &ix) {
    Vec o(ix.size());
    for (size_t k = 0; k < ix.size(); k++) o[k] = v[ix[k]];
    return o;
}

// The generated PSD function for the current VVar counters
// (DoWriteXML2m.m:148-337).
static TrOut psd_tr(const Node &root, Ctx &c) {
    // the generated function assigns its CVs on every call, so a CV may
    // depend on the VVar counters of this TR (DoWriteXML2m.m:33-39)
    for (auto &k : root.kids)
        if (k.name == "CVs")
            for (auto &a : k.attrs) c.cv[a.first] = a.second.empty() ? NAN : Expr(a.second, c).run();
    Lines all;
    Vec se;
    int sec = 0;
    for (auto &pulses : root.kids) {
        if (pulses.name != "Pulses") continue;
        sec++;
        const Params s = params_of(pulses, c);
        const double tS = s.num("tS"), tE = s.num("tE");
        if (!(tS < tE)) throw Error("SE setting is incorrect for Pulses " + std::to_string(sec) + "!");
        if (!section_on(s, c)) continue;
        Lines L;
        for (auto &grp : pulses.kids) {
            for (auto &el : grp.kids) {
                const Params p = params_of(el, c);
                if (p.text("Switch") != "on") continue;
                if (grp.name == "rf") {
                    if (p.text("AnchorTE") == "on") {
                        const std::string a = c.vtext("TEAnchor", "Middle");
                        if (a == "Start") c.num["TEAnchorTime"] = {p.num("tStart")};
                        else if (a == "Middle") c.num["TEAnchorTime"] = {(p.num("tStart") + p.num("tEnd")) / 2};
                        else if (a == "End") c.num["TEAnchorTime"] = {p.num("tEnd")};
                    }
                    RfOut r = rf_element(el.name, p, c);
                    const bool multi = c.vtext("MultiTransmit", "off") != "off";
                    if (multi || c.vctl("MasterTxCoil") == r.coil.front()) {
                        append(L.rfAmp, r.amp);
                        append(L.rfPhase, r.phase);
                        append(L.rfFreq, r.freq);
                        append(L.rfCoil, r.coil);
                        append(L.rfT, r.t);
                    }
                } else if (grp.name == "GzSS") {
                    GOut g;
                    if (el.name == "GzCartesian") g = g_cartesian_pe(p, c, true);
                    else if (el.name == "GzSelective") g = gz_selective(p);
                    else if (el.name == "GzSelective2") g = gz_selective2(p);
                    else if (!g_trapezoid(el.name, "Gz", p, c, g))
                        throw Error("GzSS element " + el.name + " is not supported");
                    append(L.Gz, g.amp);
                    append(L.GzT, g.t);
                } else if (grp.name == "GyPE") {
                    GOut g;
                    if (el.name == "GyCartesian") g = g_cartesian_pe(p, c, false);
                    else if (el.name == "GyEPI") g = gy_epi(p, c);
                    else if (el.name == "GyFSE") g = gy_fse(p, c);
                    else if (el.name == "GyRadial") g = g_radial(p, c, true);
                    else if (el.name == "GySpiral") g = g_spiral(p, c, true);
                    else if (!g_trapezoid(el.name, "Gy", p, c, g))
                        throw Error("GyPE element " + el.name + " is not supported");
                    append(L.Gy, g.amp);
                    append(L.GyT, g.t);
                } else if (grp.name == "GxR") {
                    GOut g;
                    if (el.name == "GxCartesian") g = gx_cartesian(p, c);
                    else if (el.name == "GxEPI") g = gx_epi(p, c);
                    else if (el.name == "GxFSE") g = gx_fse(p, c);
                    else if (el.name == "GxRadial") g = g_radial(p, c, false);
                    else if (el.name == "GxSpiral") g = g_spiral(p, c, false);
                    else if (!g_trapezoid(el.name, "Gx", p, c, g))
                        throw Error("GxR element " + el.name + " is not supported");
                    append(L.Gx, g.amp);
                    append(L.GxT, g.t);
                } else if (grp.name == "ADC") {
                    GOut g;
                    if (el.name == "ADCCartesian") g = adc_cartesian(p, c);
                    else if (el.name == "ADCEPI") g = adc_epi(p, c);
                    else if (el.name == "ADCFSE") g = adc_fse(p, c);
                    else if (el.name == "ADCRadial") g = adc_radial(p, c);
                    else if (el.name == "ADCSpiral") g = adc_spiral(p, c);
                    else throw Error("ADC element " + el.name + " is not supported");
                    append(L.ADC, g.amp);
                    append(L.ADCT, g.t);
                } else if (grp.name == "Ext") {
                    if (el.name != "ExtBit") throw Error("Ext element " + el.name + " is not supported");
                    GOut g = ext_bit(p, c);
                    append(L.Ext, g.amp);
                    append(L.ExtT, g.t);
                } else {
                    throw Error("unknown sequence line " + grp.name);
                }
            }
        }
        // clip to the section, rf below eps to 0, offset by tS
        const double span = tE - tS;
        auto clip = [&](std::vector<Vec *> vals, Vec &t) {
            std::vector<size_t> keep;
            for (size_t k = 0; k < t.size(); k++)
                if (!(t[k] < 0 || t[k] > span)) keep.push_back(k);
            for (Vec *v : vals) *v = take(*v, keep);
            t = take(t, keep);
            for (double &x : t) x = x + tS;
        };
        clip({&L.rfAmp, &L.rfPhase, &L.rfFreq, &L.rfCoil}, L.rfT);
        for (double &a : L.rfAmp)
            if (std::fabs(a) < kEps) a = 0;
        clip({&L.Gz}, L.GzT);
        clip({&L.Gy}, L.GyT);
        clip({&L.Gx}, L.GxT);
        clip({&L.ADC}, L.ADCT);
        clip({&L.Ext}, L.ExtT);
        append(all.rfAmp, L.rfAmp);
        append(all.rfPhase, L.rfPhase);
        append(all.rfFreq, L.rfFreq);
        append(all.rfCoil, L.rfCoil);
        append(all.rfT, L.rfT);
        append(all.Gz, L.Gz);
        append(all.GzT, L.GzT);
        append(all.Gy, L.Gy);
        append(all.GyT, L.GyT);
        append(all.Gx, L.Gx);
        append(all.GxT, L.GxT);
        append(all.ADC, L.ADC);
        append(all.ADCT, L.ADCT);
        append(all.Ext, L.Ext);
        append(all.ExtT, L.ExtT);
        se.push_back(tS);
        se.push_back(tE);
    }
    if (all.rfT.empty()) throw Error("rf sequence line can not be empty! Master Tx coil element must be used.");
    if (all.GzT.empty()) throw Error("GzSS sequence line can not be empty!");
    if (all.GyT.empty()) throw Error("GyPE sequence line can not be empty!");
    if (all.GxT.empty()) throw Error("GxR sequence line can not be empty!");
    if (all.ADCT.empty()) throw Error("ADC sequence line can not be empty!");
    if (all.ExtT.empty()) throw Error("Ext sequence line can not be empty!");
    const double t0 = *std::min_element(se.begin(), se.end()), t1 = *std::max_element(se.begin(), se.end());
    TrOut o;
    Vec ts{t0, t1};
    std::vector<int> fl{-1, -1};
    const Vec *lines[6] = {&all.rfT, &all.GzT, &all.GyT, &all.GxT, &all.ADCT, &all.ExtT};
    for (int l = 0; l < 6; l++)
        for (double t : *lines[l]) {
            ts.push_back(t);
            fl.push_back(l);
        }
    for (double &t : ts) t = t - t0;
    const auto ix = order_of(ts);
    o.ts = take(ts, ix);
    o.flags.assign(6 * ix.size(), 0.0);
    for (size_t k = 0; k < ix.size(); k++)
        if (fl[ix[k]] >= 0) o.flags[6 * k + fl[ix[k]]] = 1;
    for (size_t k = 0; k < o.ts.size(); k++)
        if (k == 0 || o.ts[k] != o.ts[k - 1]) o.uts.push_back(o.ts[k]);
    auto sorted = [&](const Vec &t, std::vector<Vec *> vals, std::vector<Vec *> outs) {
        Vec tt = t;
        for (double &x : tt) x = x - t0;
        const auto jx = order_of(tt);
        for (size_t v = 0; v < vals.size(); v++) *outs[v] = take(*vals[v], jx);
    };
    sorted(all.rfT, {&all.rfAmp, &all.rfPhase, &all.rfFreq, &all.rfCoil}, {&o.rfAmp, &o.rfPhase, &o.rfFreq, &o.rfCoil});
    sorted(all.GzT, {&all.Gz}, {&o.Gz});
    sorted(all.GyT, {&all.Gy}, {&o.Gy});
    sorted(all.GxT, {&all.Gx}, {&o.Gx});
    sorted(all.ADCT, {&all.ADC}, {&o.ADC});
    sorted(all.ExtT, {&all.Ext}, {&o.Ext});
    for (Vec *v : {&o.rfAmp, &o.rfPhase, &o.rfFreq, &o.Gz, &o.Gy, &o.Gx, &o.ADC, &o.Ext}) {
        v->front() = 0;
        v->back() = 0;
    }
    return o;
}

// DoPulseGen.m:18-117: the slice x phase loop, dummy pulses, concatenation.
static void compile(const Node &root, Ctx &c, mrs_seq &out) {
    if (root.name != "MRiLabSeq") throw Error("not a PSD: the root element is <" + root.name + ">");
    const double tr0 = c.vctl("TR"), fa0 = c.vctl("FlipAng");
    const bool dp_on = c.vtext("DP_Flag", "off") == "on";
    const double n2 = c.vctl("SecondPhNum"), n1 = c.vctl("FirstPhNum");
    if (!c.has("TEAnchorTime")) c.num["TEAnchorTime"] = {0.0};
    if (!c.has("MasterTxCoil")) c.num["MasterTxCoil"] = {1.0};
    int dp = 0;
    double s = 1, j = 1;
    c.tr_count = 0;
    while (s <= n2) {
        c.slice_count = s;
        bool restarted = false;
        while (j <= n1) {
            c.phase_count = j;
            c.tr_count = c.tr_count + 1;
            if (c.txt.count("DP_Flag")) {
                if (dp_on && dp == 0) {
                    c.num["TR"] = {c.vctl("DP_TR")};
                    c.num["FlipAng"] = {c.vctl("DP_FlipAng")};
                } else {
                    c.num["TR"] = {tr0};
                    c.num["FlipAng"] = {fa0};
                }
            }
            TrOut t = psd_tr(root, c);
            append(out.rfAmp, t.rfAmp);
            append(out.rfPhase, t.rfPhase);
            append(out.rfFreq, t.rfFreq);
            append(out.rfCoil, t.rfCoil);
            append(out.Gz, t.Gz);
            append(out.Gy, t.Gy);
            append(out.Gx, t.Gx);
            append(out.ADC, t.ADC);
            append(out.Ext, t.Ext);
            append(out.uts, t.uts);
            append(out.ts, t.ts);
            append(out.flags, t.flags);
            out.n_tr++;
            if (c.txt.count("DP_Flag") && dp_on && dp == 0) {
                for (Vec *v : {&out.Gz, &out.Gy, &out.Gx, &out.ADC}) std::fill(v->begin(), v->end(), 0.0);
                if (c.tr_count == c.vctl("DP_Num")) {
                    s = 1;
                    j = 1;
                    dp = 1;
                    restarted = true;
                    break;
                }
            }
            j = j + 1;
        }
        if (restarted) continue;
        j = 1;
        s = s + 1;
    }
    if (c.vtext("MultiTransmit", "off") == "off") std::fill(out.rfCoil.begin(), out.rfCoil.end(), 1.0);
    out.multi_transmit = c.vtext("MultiTransmit", "off") == "on";
    out.master_tx = c.vctl("MasterTxCoil");
}

}  // namespace seqgen
}  // namespace mrs

using namespace mrs;

extern "C" {

mrs_status mrs_seq_compile(const char *psd_xml, const mrs_seq_param *vctl, int32_t n_params, double gyro,
                           mrs_seq **out) {
    if (!out) return fail(MRS_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!psd_xml) return fail(MRS_ERR_ARG, "psd_xml is NULL");
    if (n_params < 0 || (n_params > 0 && !vctl)) return fail(MRS_ERR_ARG, "bad VCtl parameter list");
    if (!(gyro > 0) || !std::isfinite(gyro)) return fail(MRS_ERR_ARG, "gyro must be positive");
    try {
        seqgen::Ctx c;
        c.gyro = gyro;
        for (int32_t i = 0; i < n_params; i++) {
            const mrs_seq_param &p = vctl[i];
            if (!p.name || !*p.name) return fail(MRS_ERR_ARG, "VCtl parameter without a name");
            if (p.text) {
                c.txt[p.name] = p.text;
            } else {
                if (p.n < 1 || !p.value) return fail(MRS_ERR_ARG, std::string("VCtl.") + p.name + " has no value");
                c.num[p.name] = seqgen::Vec(p.value, p.value + p.n);
            }
        }
        const std::string text(psd_xml);
        const seqgen::Node root = seqgen::XmlReader(text).document();
        std::unique_ptr<mrs_seq> s(new mrs_seq());
        seqgen::compile(root, c, *s);
        *out = s.release();
        return MRS_OK;
    } catch (const seqgen::Error &e) {
        return fail(MRS_ERR_ARG, std::string("PSD: ") + e.what());
    } catch (const std::bad_alloc &) {
        return fail(MRS_ERR_NOMEM, "PSD: out of host memory");
    } catch (const std::exception &e) {
        return fail(MRS_ERR_ARG, std::string("PSD: ") + e.what());
    }
}

// DoMultiTransmitChk.m, DoGradChk.m and DoUpdRateChk.m on the compiled
// lines, in their order; the first failing check names itself.
static mrs_status seq_check(const mrs_seq *s, double max_grad, double max_slew, double min_upd,
                            int32_t tx_coil_num);

mrs_status mrs_seq_check(const mrs_seq *s, double max_grad, double max_slew, double min_upd, int32_t tx_coil_num) {
    try {
        return seq_check(s, max_grad, max_slew, min_upd, tx_coil_num);
    } catch (const std::bad_alloc &) {
        return fail(MRS_ERR_NOMEM, "sequence check: out of host memory");
    }
}

static mrs_status seq_check(const mrs_seq *s, double max_grad, double max_slew, double min_upd,
                            int32_t tx_coil_num) {
    if (!s) return fail(MRS_ERR_ARG, "NULL argument");
    if (s->multi_transmit && tx_coil_num > 0 && !s->rfCoil.empty()) {  // DoMultiTransmitChk.m:7-18
        const double mx = *std::max_element(s->rfCoil.begin(), s->rfCoil.end());
        if ((double)tx_coil_num < mx)
            return fail(MRS_ERR_ARG, "Multi-transmit check failed: the sequence uses more coil elements than "
                                     "the Tx coil provides");
        if (std::find(s->rfCoil.begin(), s->rfCoil.end(), s->master_tx) == s->rfCoil.end())
            return fail(MRS_ERR_ARG, "Multi-transmit check failed: the master Tx coil element is not used");
    }
    const size_t n = s->ts.size();
    auto times = [&](int row) {  // tsLine(flagsLine(row, :) ~= 0)
        std::vector<double> t;
        for (size_t k = 0; k < n; k++)
            if (s->flags[k * 6 + (size_t)row] != 0) t.push_back(s->ts[k]);
        return t;
    };
    const double gtol = std::fabs(max_grad * 0.005), stol = std::fabs(max_slew * 0.005);
    const char *gname[3] = {"GzSS", "GyP", "GxR"};
    const std::vector<double> *gline[3] = {&s->Gz, &s->Gy, &s->Gx};
    std::string why;
    for (int g = 0; g < 3 && why.empty(); g++) {  // DoGradChk.m:8-61
        const std::vector<double> t = times(g + 1), &a = *gline[g];
        if (t.size() != a.size()) return fail(MRS_ERR_ARG, std::string(gname[g]) + " line and flags disagree");
        for (size_t k = 0; k + 1 < a.size(); k++) {
            const double sr = (a[k + 1] - a[k]) / (t[k + 1] - t[k]);
            if ((std::fabs(sr) - max_slew) > stol) {
                why = std::string(gname[g]) + " exceeds maximum allowed gradient slew rate";
                break;
            }
        }
        if (!why.empty()) break;
        for (double x : a)
            if ((std::fabs(x) - max_grad) > gtol) {
                why = std::string(gname[g]) + " exceeds maximum allowed gradient amplitude";
                break;
            }
    }
    if (!why.empty()) return fail(MRS_ERR_ARG, "Gradient check failed: " + why);
    const double utol = -std::fabs(min_upd * 0.01), half_eps = 0.5 * seqgen::kEps;
    auto too_fast = [&](const std::vector<double> &t) {
        for (size_t k = 0; k + 1 < t.size(); k++) {
            const double dt = t[k + 1] - t[k];
            if (dt < half_eps) continue;  // unique's rounding threshold; TR boundaries
            if ((std::fabs(dt) - min_upd) < utol) return true;
        }
        return false;
    };
    {  // DoUpdRateChk.m:8-22, per Tx coil in ascending order
        const std::vector<double> t = times(0);
        if (t.size() != s->rfCoil.size()) return fail(MRS_ERR_ARG, "rf line and flags disagree");
        std::vector<double> ids(s->rfCoil);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        for (double id : ids) {
            std::vector<double> tc;
            for (size_t k = 0; k < t.size(); k++)
                if (s->rfCoil[k] == id) tc.push_back(t[k]);
            if (too_fast(tc)) {
                char b[96];
                std::snprintf(b, sizeof b, "rf%g sequence line exceeds minimum update rate", id);
                return fail(MRS_ERR_ARG, std::string("Update rate check failed: ") + b);
            }
        }
    }
    for (int g = 0; g < 3; g++)  // DoUpdRateChk.m:23-50
        if (too_fast(times(g + 1)))
            return fail(MRS_ERR_ARG, std::string("Update rate check failed: ") + gname[g] +
                                         " sequence line exceeds minimum update rate");
    return MRS_OK;
}

mrs_status mrs_seq_counts(const mrs_seq *s, int64_t counts[9]) {
    if (!s || !counts) return fail(MRS_ERR_ARG, "NULL argument");
    counts[0] = (int64_t)s->ts.size();
    counts[1] = (int64_t)s->uts.size();
    counts[2] = (int64_t)s->rfAmp.size();
    counts[3] = (int64_t)s->Gz.size();
    counts[4] = (int64_t)s->Gy.size();
    counts[5] = (int64_t)s->Gx.size();
    counts[6] = (int64_t)s->ADC.size();
    counts[7] = (int64_t)s->Ext.size();
    counts[8] = s->n_tr;
    return MRS_OK;
}

mrs_status mrs_seq_copy(const mrs_seq *s, const mrs_seq_lines *dst) {
    if (!s || !dst) return fail(MRS_ERR_ARG, "NULL argument");
    const std::pair<const std::vector<double> *, double *> pairs[] = {
        {&s->uts, dst->utsLine},     {&s->ts, dst->tsLine},         {&s->rfAmp, dst->rfAmpLine},
        {&s->rfPhase, dst->rfPhaseLine}, {&s->rfFreq, dst->rfFreqLine}, {&s->rfCoil, dst->rfCoilLine},
        {&s->Gz, dst->GzAmpLine},    {&s->Gy, dst->GyAmpLine},      {&s->Gx, dst->GxAmpLine},
        {&s->ADC, dst->ADCLine},     {&s->Ext, dst->ExtLine},       {&s->flags, dst->flagsLine}};
    for (auto &p : pairs) {
        if (!p.second && !p.first->empty()) return fail(MRS_ERR_ARG, "a destination line is NULL");
        if (!p.first->empty()) std::memcpy(p.second, p.first->data(), p.first->size() * sizeof(double));
    }
    return MRS_OK;
}

void mrs_seq_destroy(mrs_seq *s) { delete s; }

}  // extern "C"
