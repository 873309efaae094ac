// params.hip — reads a parameters.h case description at run time (host code only).
//
// The reference configures everything with compile-time macros (parameters.h:1-55).  This reader
// evaluates the same file's #define lines with C arithmetic rules (integer literals stay
// integers, int/int divides with truncation, mixing with a double promotes), so a case file that
// compiles the reference describes the same case here.  Macros that have no rbtp_case field
// (LXY, LXYZ, NPOP, BLOCK_*) are evaluated and ignored; unknown identifiers are an error.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "../../include/rbtp.h"

namespace rbtp {
int report(int code, const char *msg);
}

namespace {

struct Val {
    double d;
    long long i;
    bool is_int;
};

Val mk_int(long long v) { return Val{(double)v, v, true}; }
Val mk_dbl(double v) { return Val{v, 0, false}; }

struct Parser {
    const char *p;
    const std::map<std::string, Val> &defs;
    std::string err;

    void ws() {
        while (*p && isspace((unsigned char)*p)) p++;
    }
    bool fail(const std::string &m) {
        if (err.empty()) err = m;
        return false;
    }
    bool primary(Val &v) {
        ws();
        if (*p == '(') {
            p++;
            if (!expr(v)) return false;
            ws();
            if (*p != ')') return fail("missing ')'");
            p++;
            return true;
        }
        if (*p == '-' || *p == '+') {
            const char s = *p++;
            if (!primary(v)) return false;
            if (s == '-') {
                v.d = -v.d;
                v.i = -v.i;
            }
            return true;
        }
        if (isdigit((unsigned char)*p) || *p == '.') {
            const char *start = p;
            bool dbl = false;
            while (isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' ||
                   ((*p == '+' || *p == '-') && (p[-1] == 'e' || p[-1] == 'E'))) {
                if (*p == '.' || *p == 'e' || *p == 'E') dbl = true;
                p++;
            }
            std::string tok(start, p);
            while (*p == 'f' || *p == 'F' || *p == 'l' || *p == 'L' || *p == 'u' || *p == 'U') {
                if (*p == 'f' || *p == 'F') dbl = true;
                p++;
            }
            if (dbl) v = mk_dbl(strtod(tok.c_str(), nullptr));
            else v = mk_int(strtoll(tok.c_str(), nullptr, 10));
            return true;
        }
        if (isalpha((unsigned char)*p) || *p == '_') {
            const char *start = p;
            while (isalnum((unsigned char)*p) || *p == '_') p++;
            std::string id(start, p);
            auto it = defs.find(id);
            if (it == defs.end()) return fail("unknown identifier " + id);
            v = it->second;
            return true;
        }
        return fail(std::string("unexpected '") + (*p ? std::string(1, *p) : "end") + "'");
    }
    static Val binop(char op, const Val &a, const Val &b) {
        if (a.is_int && b.is_int) {
            switch (op) {
                case '+': return mk_int(a.i + b.i);
                case '-': return mk_int(a.i - b.i);
                case '*': return mk_int(a.i * b.i);
                default: return mk_int(b.i == 0 ? 0 : a.i / b.i);
            }
        }
        switch (op) {
            case '+': return mk_dbl(a.d + b.d);
            case '-': return mk_dbl(a.d - b.d);
            case '*': return mk_dbl(a.d * b.d);
            default: return mk_dbl(a.d / b.d);
        }
    }
    bool term(Val &v) {
        if (!primary(v)) return false;
        for (;;) {
            ws();
            if (*p != '*' && *p != '/') return true;
            const char op = *p++;
            Val r;
            if (!primary(r)) return false;
            v = binop(op, v, r);
        }
    }
    bool expr(Val &v) {
        if (!term(v)) return false;
        for (;;) {
            ws();
            if (*p != '+' && *p != '-') return true;
            const char op = *p++;
            Val r;
            if (!term(r)) return false;
            v = binop(op, v, r);
        }
    }
};

std::string strip_comments(const std::string &text) {
    std::string out;
    for (size_t i = 0; i < text.size();) {
        if (text.compare(i, 2, "//") == 0) {
            while (i < text.size() && text[i] != '\n') i++;
        } else if (text.compare(i, 2, "/*") == 0) {
            const size_t e = text.find("*/", i + 2);
            i = (e == std::string::npos) ? text.size() : e + 2;
        } else {
            out += text[i++];
        }
    }
    return out;
}

}  // namespace

extern "C" int rbtp_case_from_parameters_h(const char *path, rbtp_case *out) {
    if (!path || !out) return rbtp::report(RBTP_E_ARG, "rbtp_case_from_parameters_h: null argument");
    FILE *fp = fopen(path, "r");
    if (!fp) return rbtp::report(RBTP_E_IO, (std::string("cannot open ") + path).c_str());
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, got);
    fclose(fp);
    text = strip_comments(text);
    std::map<std::string, Val> defs;
    size_t pos = 0;
    int lineno = 0;
    while (pos < text.size()) {
        size_t e = text.find('\n', pos);
        if (e == std::string::npos) e = text.size();
        std::string line = text.substr(pos, e - pos);
        pos = e + 1;
        lineno++;
        const char *q = line.c_str();
        while (isspace((unsigned char)*q)) q++;
        if (strncmp(q, "#define", 7) != 0) continue;
        q += 7;
        while (isspace((unsigned char)*q)) q++;
        const char *ns = q;
        while (isalnum((unsigned char)*q) || *q == '_') q++;
        std::string name(ns, q);
        while (isspace((unsigned char)*q)) q++;
        if (*q == '\0') continue;  // include guard
        Parser P{q, defs, ""};
        Val v;
        if (!P.expr(v)) {
            char msg[512];
            snprintf(msg, sizeof(msg), "%s:%d: %s: %s", path, lineno, name.c_str(), P.err.c_str());
            return rbtp::report(RBTP_E_IO, msg);
        }
        P.ws();
        if (*P.p != '\0') {
            char msg[512];
            snprintf(msg, sizeof(msg), "%s:%d: trailing text after %s", path, lineno, name.c_str());
            return rbtp::report(RBTP_E_IO, msg);
        }
        defs[name] = v;
    }
    rbtp_case c;
    rbtp_case_default(&c);
    auto I = [&](const char *k, int &dst) {
        auto it = defs.find(k);
        if (it != defs.end()) dst = it->second.is_int ? (int)it->second.i : (int)it->second.d;
    };
    auto D = [&](const char *k, double &dst) {
        auto it = defs.find(k);
        if (it != defs.end()) dst = it->second.d;
    };
    I("CONTINUOUS_STEPS", c.continuous_steps);
    I("LX", c.lx); I("LY", c.ly); I("LZ", c.lz);
    I("NEND", c.nend); I("NDIAG", c.ndiag); I("NFLOWOUT", c.nflowout); I("NNUOUT", c.nnuout);
    I("NOUT_P", c.nout_p);
    D("RAYLEIGH", c.rayleigh); D("PRANDTL", c.prandtl); D("T_HOT", c.t_hot); D("T_COLD", c.t_cold);
    D("T_REF", c.t_ref); D("BETA", c.beta); D("VISC", c.visc); D("RHO0", c.rho0);
    I("ACTIVATE_PARTICLES", c.activate_particles);
    D("PARTICLE_FRACTION", c.particle_fraction); D("PARTICLE_DIAM_FRAC", c.particle_diam_frac);
    D("CP_PARTICLE_RATIO", c.cp_particle_ratio); D("CP_FLUID", c.cp_fluid);
    D("RHO_PARTICLE_RATIO", c.rho_particle_ratio); D("REPFORCE_SCALE", c.repforce_scale);
    D("PARTICLE_INIT_TEMP", c.particle_init_temp);
    I("PARTICLE_RAND_SEED", c.particle_rand_seed);
    D("REP_MIN_GAP", c.rep_min_gap); D("REP_MIN_GAP_W", c.rep_min_gap_w);
    D("REP_STF0", c.rep_stf0); D("REP_STF1", c.rep_stf1);
    D("REP_STF0_W", c.rep_stf0_w); D("REP_STF1_W", c.rep_stf1_w);
    I("NPART_OVERRIDE", c.npart_override);
    I("PRECISION", c.precision);
    *out = c;
    return RBTP_OK;
}
