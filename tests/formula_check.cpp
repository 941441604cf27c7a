// formula_check.cpp -- host harness for the MathTree compiler: random formula
// trees are linearised the way a VXA formula is read (reversed breadth-first
// order, MT_END last), evaluated by the FIFO evaluator (mt_eval) and, after
// mt_postfix, by the bounded-stack evaluator the device runs (formula_eval).
// Built by tests/test_formula.py with g++ together with vx3mi_formula.cpp.
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "vx3mi_internal.h"

namespace vx3mi {
void set_error(const std::string &) {}
}  // namespace vx3mi

using namespace vx3mi;

namespace {

struct TNode {
    MtToken tk;
    std::vector<int> kids;
};

int arity(int op) {
    if (op == MT_CONST || op == MT_E || op == MT_PI || op == MT_VAR) return 0;
    if ((op >= MT_SQRT && op <= MT_NOT) || op == MT_NORMALCDF) return 1;
    return 2;
}

int grow(std::vector<TNode> &t, std::mt19937_64 &rng, int depth) {
    static const int ops[] = {MT_ADD, MT_SUB, MT_MUL, MT_DIV, MT_POW, MT_SQRT, MT_SIN, MT_COS, MT_TAN, MT_ATAN,
                              MT_LOG, MT_INT, MT_ABS, MT_NOT, MT_GT, MT_LT, MT_AND, MT_OR, MT_NORMALCDF};
    std::uniform_real_distribution<double> u(0, 1);
    TNode n;
    if (depth <= 0 || u(rng) < 0.25) {
        const double r = u(rng);
        if (r < 0.45) n.tk = MtToken{MT_CONST, std::floor(u(rng) * 200 - 100) / 8};
        else if (r < 0.9) n.tk = MtToken{MT_VAR, (double)(int)(u(rng) * MV_COUNT)};
        else n.tk = MtToken{r < 0.95 ? MT_E : MT_PI, 0.0};
    } else {
        n.tk = MtToken{ops[(int)(u(rng) * (sizeof(ops) / sizeof(ops[0])))], 0.0};
    }
    const int id = (int)t.size();
    t.push_back(n);
    for (int k = 0; k < arity(n.tk.op); k++) {
        const int c = grow(t, rng, depth - 1);
        t[id].kids.push_back(c);
    }
    return id;
}

}  // namespace

// n random formulas of up to `depth` levels; returns the number whose two
// evaluations differ in any bit (NaN equals NaN), -1 on a compile failure;
// *deepest receives the largest postfix stack depth met
extern "C" int check_random(uint64_t seed, int n, int depth, int *skipped) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-3, 3);
    int bad = 0;
    *skipped = 0;
    for (int i = 0; i < n; i++) {
        std::vector<TNode> t;
        const int root = grow(t, rng, depth);
        // reversed BFS of the subtree below the formula element (the root
        // is the formula's only child), then MT_END
        std::vector<int> bfs{root};
        for (size_t h = 0; h < bfs.size(); h++)
            for (int c : t[bfs[h]].kids) bfs.push_back(c);
        std::vector<MtToken> fifo;
        for (auto it = bfs.rbegin(); it != bfs.rend(); ++it) fifo.push_back(t[*it].tk);
        fifo.push_back(MtToken{MT_END, 0.0});
        std::vector<MtToken> post;
        if (!mt_postfix(fifo, post)) { (*skipped)++; continue; }  // deeper than the device stack
        for (int trial = 0; trial < 4; trial++) {
            double vars[MV_COUNT];
            for (double &v : vars) v = u(rng);
            double a = 0;
            if (!mt_eval(fifo, vars, &a)) return -1;
            const double b = formula_eval(post.data(), (int)post.size(), vars);
            if (std::memcmp(&a, &b, sizeof a) != 0 && !(a != a && b != b)) bad++;
        }
    }
    return bad;
}
