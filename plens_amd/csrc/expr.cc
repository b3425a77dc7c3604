// Compiler of the varying-inflow expressions (deal.II FunctionParser over muParser): recursive
// descent in muParser's operator precedence, from lowest to highest:
//   c ? a : b    (right associative)
//   ||           &&           == != < > <= >=
//   + -          * /          unary - +        ^ (right associative, binds tighter than unary -)
// functions: sin cos tan asin acos atan atan2 sinh cosh tanh asinh acosh atanh log2 log10 log ln
// exp sqrt sign rint abs, min max sum avg (one or more arguments), and deal.II's additions if
// (= c ? a : b), pow, floor, ceil, int, cot, csc, sec, erfc; constants _pi, _e; variables x y z t.
#include "expr.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Compiler {
    const char* s;
    size_t i = 0;
    ExprCode* out;
    int depth = 0, max_depth = 0;
    std::string err;

    void ws() {
        while (s[i] && std::isspace((unsigned char)s[i])) ++i;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m + " at position " + std::to_string(i);
        return false;
    }
    bool emit(ExprOp o, double k, int pops, int pushes) {
        if (out->n >= kExprMaxOps) return fail("expression too long");
        out->op[out->n] = o;
        out->k[out->n] = k;
        ++out->n;
        depth += pushes - pops;
        if (depth > max_depth) max_depth = depth;
        if (max_depth > kExprMaxStack) return fail("expression nests too deeply");
        return true;
    }
    bool match(const char* tok) {
        ws();
        const size_t n = std::strlen(tok);
        if (std::strncmp(s + i, tok, n) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    bool ternary() {
        if (!lor()) return false;
        ws();
        if (s[i] == '?') {
            ++i;
            if (!ternary()) return false;
            if (!match(":")) return fail("':' expected");
            if (!ternary()) return false;
            return emit(EOP_SELECT, 0, 3, 1);
        }
        return true;
    }
    bool lor() {
        if (!land()) return false;
        while (match("||"))
            if (!land() || !emit(EOP_OR, 0, 2, 1)) return false;
        return true;
    }
    bool land() {
        if (!cmp()) return false;
        while (match("&&"))
            if (!cmp() || !emit(EOP_AND, 0, 2, 1)) return false;
        return true;
    }
    bool cmp() {
        if (!add()) return false;
        for (;;) {
            ExprOp o;
            if (match("<=")) o = EOP_LE;
            else if (match(">=")) o = EOP_GE;
            else if (match("==")) o = EOP_EQ;
            else if (match("!=")) o = EOP_NE;
            else if (match("<")) o = EOP_LT;
            else if (match(">")) o = EOP_GT;
            else return true;
            if (!add() || !emit(o, 0, 2, 1)) return false;
        }
    }
    bool add() {
        if (!mul()) return false;
        for (;;) {
            ws();
            if (s[i] != '+' && s[i] != '-') return true;
            const ExprOp o = s[i] == '+' ? EOP_ADD : EOP_SUB;
            ++i;
            if (!mul() || !emit(o, 0, 2, 1)) return false;
        }
    }
    bool mul() {
        if (!unary()) return false;
        for (;;) {
            ws();
            if (s[i] != '*' && s[i] != '/') return true;
            const ExprOp o = s[i] == '*' ? EOP_MUL : EOP_DIV;
            ++i;
            if (!unary() || !emit(o, 0, 2, 1)) return false;
        }
    }
    bool unary() {
        ws();
        if (s[i] == '-') {
            ++i;
            return unary() && emit(EOP_NEG, 0, 1, 1);
        }
        if (s[i] == '+') {
            ++i;
            return unary();
        }
        return power();
    }
    bool power() {
        if (!primary()) return false;
        ws();
        if (s[i] == '^') {
            ++i;
            if (!unary()) return false;  // right associative; the exponent may carry a sign
            return emit(EOP_POW, 0, 2, 1);
        }
        return true;
    }
    bool primary() {
        ws();
        const char ch = s[i];
        if (ch == '(') {
            ++i;
            if (!ternary()) return false;
            if (!match(")")) return fail("')' expected");
            return true;
        }
        if (std::isdigit((unsigned char)ch) || ch == '.') {
            char* end = nullptr;
            const double v = std::strtod(s + i, &end);
            if (end == s + i) return fail("bad number");
            i = end - s;
            return emit(EOP_CONST, v, 0, 1);
        }
        if (std::isalpha((unsigned char)ch) || ch == '_') {
            const size_t b = i;
            while (std::isalnum((unsigned char)s[i]) || s[i] == '_') ++i;
            const std::string id(s + b, i - b);
            ws();
            if (s[i] == '(') return call(id);
            if (id == "x") return emit(EOP_X, 0, 0, 1);
            if (id == "y") return emit(EOP_Y, 0, 0, 1);
            if (id == "z") return emit(EOP_Z, 0, 0, 1);
            if (id == "t") return emit(EOP_T, 0, 0, 1);
            if (id == "_pi") return emit(EOP_CONST, 3.141592653589793238462643, 0, 1);
            if (id == "_e") return emit(EOP_CONST, 2.718281828459045235360287, 0, 1);
            i = b;
            return fail("unknown name '" + id + "'");
        }
        return fail("unexpected character");
    }
    bool call(const std::string& f) {
        ++i;  // '('
        int na = 0;
        ws();
        if (s[i] != ')') {
            for (;;) {
                if (!ternary()) return false;
                ++na;
                if (match(",")) continue;
                break;
            }
        }
        if (!match(")")) return fail("')' expected after the arguments of " + f);
        struct F { const char* name; ExprOp op; };
        static const F unary_fns[] = {
            {"sin", EOP_SIN}, {"cos", EOP_COS}, {"tan", EOP_TAN}, {"asin", EOP_ASIN}, {"acos", EOP_ACOS},
            {"atan", EOP_ATAN}, {"sinh", EOP_SINH}, {"cosh", EOP_COSH}, {"tanh", EOP_TANH},
            {"asinh", EOP_ASINH}, {"acosh", EOP_ACOSH}, {"atanh", EOP_ATANH}, {"log2", EOP_LOG2},
            {"log10", EOP_LOG10}, {"log", EOP_LN}, {"ln", EOP_LN}, {"exp", EOP_EXP}, {"sqrt", EOP_SQRT},
            {"sign", EOP_SIGN}, {"rint", EOP_RINT}, {"abs", EOP_ABS}, {"floor", EOP_FLOOR},
            {"ceil", EOP_CEIL}, {"int", EOP_INT}, {"cot", EOP_COT}, {"csc", EOP_CSC}, {"sec", EOP_SEC},
            {"erfc", EOP_ERFC}};
        for (const F& u : unary_fns)
            if (f == u.name) {
                if (na != 1) return fail(f + " takes one argument");
                return emit(u.op, 0, 1, 1);
            }
        if (f == "pow" || f == "atan2") {
            if (na != 2) return fail(f + " takes two arguments");
            return emit(f == "pow" ? EOP_POW : EOP_ATAN2, 0, 2, 1);
        }
        if (f == "if") {
            if (na != 3) return fail("if takes three arguments");
            return emit(EOP_SELECT, 0, 3, 1);
        }
        const ExprOp vo = f == "min" ? EOP_MIN : f == "max" ? EOP_MAX : f == "sum" ? EOP_SUM : f == "avg" ? EOP_AVG
                                                                                                          : EOP_CONST;
        if (vo != EOP_CONST) {
            if (na < 1) return fail(f + " needs an argument");
            return emit(vo, na, na, 1);
        }
        return fail("unknown function '" + f + "'");
    }
};

}  // namespace

bool expr_compile(const char* src, ExprCode* out, char* err, int err_len) {
    if (!src || !out) return false;
    *out = ExprCode{};
    Compiler c{src, 0, out};
    bool ok = c.ternary();
    if (ok) {
        c.ws();
        if (src[c.i] != '\0') ok = c.fail("unexpected text");
    }
    if (ok && c.depth != 1) ok = c.fail("malformed expression");
    if (!ok && err && err_len > 0) std::snprintf(err, err_len, "%s", c.err.c_str());
    return ok;
}

bool expr_compile_components(const char* src, int n_components, ExprCode* out, char* err, int err_len) {
    if (!src) return false;
    std::vector<std::string> parts;
    std::string cur;
    for (const char* p = src; *p; ++p) {
        if (*p == ';') {
            parts.push_back(cur);
            cur.clear();
        } else {
            cur += *p;
        }
    }
    parts.push_back(cur);
    if ((int)parts.size() != n_components) {
        if (err && err_len > 0)
            std::snprintf(err, err_len, "%d components expected, %d given", n_components, (int)parts.size());
        return false;
    }
    for (int k = 0; k < n_components; ++k)
        if (!expr_compile(parts[k].c_str(), &out[k], err, err_len)) return false;
    return true;
}

extern "C" int plens_expr_eval(const char* expr, int32_t n_components, double x, double y, double z, double t,
                               double* values) {
    if (!expr || !values || n_components < 1 || n_components > 8) return -1;
    ExprCode code[8];
    if (!expr_compile_components(expr, n_components, code, nullptr, 0)) return -1;
    for (int k = 0; k < n_components; ++k) values[k] = expr_eval(code[k], x, y, z, t);
    return 0;
}
