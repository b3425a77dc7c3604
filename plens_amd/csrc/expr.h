// Expressions of the "varying inflow" BC (IBCs/varying_inflow.cc): deal.II FunctionParser /
// muParser syntax in the variables x, y, z, t, compiled on the host into a postfix program that
// host and device evaluate with the same routine.
#ifndef PLENS_EXPR_H
#define PLENS_EXPR_H
#include <stdint.h>

#include <cmath>

#ifdef __HIPCC__
#define PLENS_HD __host__ __device__
#else
#define PLENS_HD
#endif

constexpr int kExprMaxOps = 128;   // postfix program length
constexpr int kExprMaxStack = 24;  // evaluation stack

enum ExprOp : uint8_t {
    EOP_CONST, EOP_X, EOP_Y, EOP_Z, EOP_T,
    EOP_ADD, EOP_SUB, EOP_MUL, EOP_DIV, EOP_POW, EOP_NEG,
    EOP_LT, EOP_GT, EOP_LE, EOP_GE, EOP_EQ, EOP_NE, EOP_AND, EOP_OR, EOP_SELECT,  // c ? a : b
    EOP_SIN, EOP_COS, EOP_TAN, EOP_ASIN, EOP_ACOS, EOP_ATAN, EOP_SINH, EOP_COSH, EOP_TANH,
    EOP_ASINH, EOP_ACOSH, EOP_ATANH, EOP_LOG2, EOP_LOG10, EOP_LN, EOP_EXP, EOP_SQRT, EOP_SIGN,
    EOP_RINT, EOP_ABS, EOP_FLOOR, EOP_CEIL, EOP_INT, EOP_COT, EOP_CSC, EOP_SEC, EOP_ERFC,
    EOP_ATAN2,
    EOP_MIN, EOP_MAX, EOP_SUM, EOP_AVG  // variadic: argument count in k[]
};

struct ExprCode {
    int32_t n = 0;
    uint8_t op[kExprMaxOps];
    double k[kExprMaxOps];
};

PLENS_HD inline double expr_eval(const ExprCode& e, double x, double y, double z, double t) {
    double s[kExprMaxStack];
    int sp = 0;
    for (int i = 0; i < e.n; ++i) {
        const int o = e.op[i];
        switch (o) {
            case EOP_CONST: s[sp++] = e.k[i]; break;
            case EOP_X: s[sp++] = x; break;
            case EOP_Y: s[sp++] = y; break;
            case EOP_Z: s[sp++] = z; break;
            case EOP_T: s[sp++] = t; break;
            case EOP_NEG: s[sp - 1] = -s[sp - 1]; break;
            case EOP_SELECT: {
                const double c = s[sp - 3], a = s[sp - 2], b = s[sp - 1];
                sp -= 2;
                s[sp - 1] = c != 0.0 ? a : b;
                break;
            }
            case EOP_MIN: case EOP_MAX: case EOP_SUM: case EOP_AVG: {
                const int na = (int)e.k[i];
                double r = s[sp - na];
                for (int j = sp - na + 1; j < sp; ++j) {
                    if (o == EOP_MIN) r = s[j] < r ? s[j] : r;
                    else if (o == EOP_MAX) r = s[j] > r ? s[j] : r;
                    else r += s[j];
                }
                if (o == EOP_AVG) r /= na;
                sp -= na;
                s[sp++] = r;
                break;
            }
            default:
                if (o >= EOP_ADD && o <= EOP_OR && o != EOP_NEG) {  // binary
                    const double a = s[sp - 2], b = s[sp - 1];
                    double r = 0;
                    switch (o) {
                        case EOP_ADD: r = a + b; break;
                        case EOP_SUB: r = a - b; break;
                        case EOP_MUL: r = a * b; break;
                        case EOP_DIV: r = a / b; break;
                        case EOP_POW: r = pow(a, b); break;
                        case EOP_LT: r = a < b; break;
                        case EOP_GT: r = a > b; break;
                        case EOP_LE: r = a <= b; break;
                        case EOP_GE: r = a >= b; break;
                        case EOP_EQ: r = a == b; break;
                        case EOP_NE: r = a != b; break;
                        case EOP_AND: r = (a != 0.0) && (b != 0.0); break;
                        default: r = (a != 0.0) || (b != 0.0); break;
                    }
                    s[--sp - 1] = r;
                } else if (o == EOP_ATAN2) {
                    const double a = s[sp - 2], b = s[sp - 1];
                    s[--sp - 1] = atan2(a, b);
                } else {  // unary function
                    const double a = s[sp - 1];
                    double r = 0;
                    switch (o) {
                        case EOP_SIN: r = sin(a); break;
                        case EOP_COS: r = cos(a); break;
                        case EOP_TAN: r = tan(a); break;
                        case EOP_ASIN: r = asin(a); break;
                        case EOP_ACOS: r = acos(a); break;
                        case EOP_ATAN: r = atan(a); break;
                        case EOP_SINH: r = sinh(a); break;
                        case EOP_COSH: r = cosh(a); break;
                        case EOP_TANH: r = tanh(a); break;
                        case EOP_ASINH: r = asinh(a); break;
                        case EOP_ACOSH: r = acosh(a); break;
                        case EOP_ATANH: r = atanh(a); break;
                        case EOP_LOG2: r = log2(a); break;
                        case EOP_LOG10: r = log10(a); break;
                        case EOP_LN: r = log(a); break;
                        case EOP_EXP: r = exp(a); break;
                        case EOP_SQRT: r = sqrt(a); break;
                        case EOP_SIGN: r = (a > 0) ? 1.0 : ((a < 0) ? -1.0 : 0.0); break;
                        case EOP_RINT: r = rint(a); break;
                        case EOP_ABS: r = fabs(a); break;
                        case EOP_FLOOR: r = floor(a); break;
                        case EOP_CEIL: r = ceil(a); break;
                        case EOP_INT: r = (double)(long long)a; break;
                        case EOP_COT: r = 1.0 / tan(a); break;
                        case EOP_CSC: r = 1.0 / sin(a); break;
                        case EOP_SEC: r = 1.0 / cos(a); break;
                        default: r = erfc(a); break;
                    }
                    s[sp - 1] = r;
                }
        }
    }
    return sp == 1 ? s[0] : NAN;
}

// Compile one expression (no ';'); false and a message on a syntax error or a program that
// exceeds the length or stack limits.
bool expr_compile(const char* src, ExprCode* out, char* err, int err_len);
// Split "a; b; c" into components and compile each (n_components of them, else false).
bool expr_compile_components(const char* src, int n_components, ExprCode* out, char* err, int err_len);

#endif
