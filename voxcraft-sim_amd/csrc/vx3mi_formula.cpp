// vx3mi_formula.cpp -- evaluation of VXA MathTree formulas (stop condition,
// fitness) on the host.
//
// Semantics of VX3_MathTree::eval (src/Utils/VX3_MathTree.h:50-190): values
// are appended in token order; a unary operator consumes the oldest
// unconsumed value, a binary operator the two oldest (left operand = the
// second of them); MT_END returns the oldest unconsumed value.  Unlike the
// reference, a token list whose MT_END finds no unconsumed value is reported
// as malformed instead of returning an uninitialised slot.  sin and cos are
// the shared vxm_ functions (the device force-field evaluator uses the same).
//
// mt_postfix turns the FIFO token list into a postfix program for the
// bounded-stack evaluator formula_eval (vx3mi_physics.h) used on the device.
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
            case MT_SIN: r = vxm_sin(a); break;
            case MT_COS: r = vxm_cos(a); break;
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

bool mt_postfix(const std::vector<MtToken> &f, std::vector<MtToken> &post) {
    // run the FIFO machine on expression nodes instead of values, then emit
    // the tree of the returned value in postfix order
    struct Node {
        MtToken tk;
        int a, b;  // operand nodes (-1 none); a is the left operand
    };
    std::vector<Node> nodes;
    std::vector<int> val;
    size_t head = 0;
    int result = -1;
    for (const MtToken &tk : f) {
        if (tk.op == MT_END) {
            if (head >= val.size()) return false;
            result = val[head];
            break;
        }
        Node nd{tk, -1, -1};
        if (tk.op == MT_CONST || tk.op == MT_E || tk.op == MT_PI || tk.op == MT_VAR) {
        } else if ((tk.op >= MT_SQRT && tk.op <= MT_NOT) || tk.op == MT_NORMALCDF) {
            if (head + 1 > val.size()) return false;
            nd.a = val[head++];
        } else if ((tk.op >= MT_ADD && tk.op <= MT_POW) || (tk.op >= MT_GT && tk.op <= MT_OR)) {
            if (head + 2 > val.size()) return false;
            nd.b = val[head];
            nd.a = val[head + 1];
            head += 2;
        } else {
            return false;
        }
        val.push_back((int)nodes.size());
        nodes.push_back(nd);
    }
    if (result < 0) return false;
    post.clear();
    int depth = 0, maxDepth = 0;
    std::vector<std::pair<int, int>> todo{{result, 0}};  // (node, operands emitted)
    while (!todo.empty()) {
        auto &t = todo.back();
        const Node &nd = nodes[t.first];
        if (t.second == 0 && nd.a >= 0) { t.second = 1; todo.push_back({nd.a, 0}); continue; }
        if (t.second <= 1 && nd.b >= 0) { t.second = 2; todo.push_back({nd.b, 0}); continue; }
        post.push_back(nd.tk);
        depth += (nd.a < 0 ? 1 : 0) - (nd.b >= 0 ? 1 : 0);
        maxDepth = depth > maxDepth ? depth : maxDepth;
        todo.pop_back();
    }
    return maxDepth <= kMaxFormulaStack;
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
