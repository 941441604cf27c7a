// vx3mi_formula.cpp -- evaluation of VXA MathTree formulas (stop condition,
// fitness) on the host.
//
// Semantics of VX3_MathTree::eval (src/Utils/VX3_MathTree.h:50-190): values
// are appended in token order; a unary operator consumes the oldest
// unconsumed value, a binary operator the two oldest (left operand = the
// second of them); MT_END returns the oldest unconsumed value.  Unlike the
// reference, a token list whose MT_END finds no unconsumed value is reported
// as malformed instead of returning an uninitialised slot.
#include <cmath>
#include <vector>

#include "vx3mi_internal.h"

namespace vx3mi {

bool mt_eval(const std::vector<MtToken> &f, const double vars[MV_COUNT], double *out) {
    std::vector<double> val(f.size() + 1);
    size_t head = 0, tail = 0;  // oldest unconsumed value, next free slot
    for (const MtToken &tk : f) {
        const size_t avail = tail - head;
        auto need = [&](size_t n) { return avail >= n; };
        double r;
        switch (tk.op) {
        case MT_END:
            if (!need(1)) return false;
            *out = val[head];
            return true;
        case MT_CONST: r = tk.value; break;
        case MT_E: r = 2.71828182845904523536; break;
        case MT_PI: r = 3.14159265358979323846; break;
        case MT_VAR: {
            const int v = (int)tk.value;
            if (v < 0 || v >= MV_COUNT) return false;
            r = vars[v];
            break;
        }
        case MT_SQRT: case MT_SIN: case MT_COS: case MT_TAN: case MT_ATAN: case MT_LOG:
        case MT_INT: case MT_ABS: case MT_NOT: case MT_NORMALCDF: {
            if (!need(1)) return false;
            const double a = val[head++];
            switch (tk.op) {
            case MT_SQRT: r = std::sqrt(a); break;
            case MT_SIN: r = std::sin(a); break;
            case MT_COS: r = std::cos(a); break;
            case MT_TAN: r = std::tan(a); break;
            case MT_ATAN: r = std::atan(a); break;
            case MT_LOG: r = std::log(a); break;
            case MT_INT: r = std::rint(a); break;
            case MT_ABS: r = std::fabs(a); break;
            case MT_NOT: r = (a == 0.0) ? 1.0 : 0.0; break;
            default: r = 0.5 * std::erfc(-a / std::sqrt(2.0)); break;  // standard normal CDF
            }
            break;
        }
        case MT_ADD: case MT_SUB: case MT_MUL: case MT_DIV: case MT_POW:
        case MT_GT: case MT_LT: case MT_AND: case MT_OR: {
            if (!need(2)) return false;
            const double b = val[head], a = val[head + 1];  // a op b
            head += 2;
            switch (tk.op) {
            case MT_ADD: r = a + b; break;
            case MT_SUB: r = a - b; break;
            case MT_MUL: r = a * b; break;
            case MT_DIV: r = a / b; break;
            case MT_POW: r = std::pow(a, b); break;
            case MT_GT: r = a > b; break;
            case MT_LT: r = a < b; break;
            case MT_AND: r = (a != 0.0) && (b != 0.0); break;
            default: r = (a != 0.0) || (b != 0.0); break;
            }
            break;
        }
        default:
            return false;
        }
        val[tail++] = r;
    }
    return false;  // no MT_END
}

unsigned mt_vars_used(const std::vector<MtToken> &f) {
    unsigned m = 0;
    for (const MtToken &tk : f)
        if (tk.op == MT_VAR && tk.value >= 0 && tk.value < MV_COUNT) m |= 1u << (int)tk.value;
    return m;
}

}  // namespace vx3mi

using namespace vx3mi;

extern "C" int vx3mi_model_eval_formula(const vx3mi_model *m, int which, const double *vars, double *out) {
    if (!m || !vars || !out || (which != VX3MI_FORMULA_STOP && which != VX3MI_FORMULA_FITNESS)) {
        set_error("vx3mi_model_eval_formula: invalid argument");
        return VX3MI_E_ARG;
    }
    const std::vector<MtToken> &f = which == VX3MI_FORMULA_STOP ? m->m.stopFormula : m->m.fitnessFormula;
    if (f.empty()) {
        set_error("the model has no such formula");
        return VX3MI_E_ARG;
    }
    if (!mt_eval(f, vars, out)) {
        set_error("malformed formula");
        return VX3MI_E_PARSE;
    }
    return VX3MI_OK;
}
