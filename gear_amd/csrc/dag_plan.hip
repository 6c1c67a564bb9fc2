// Plan builder for expression programs: compiles every row of every program
// into the tape of dag_tape.h and uploads it to the current device.
//
// The backward-pass bookkeeping resolved here follows the reference's
// backward_general queue (auto_LiRPA/backward_bound.py:989-1125): a node is
// visited once all of its consumers have pushed their A into it, ancestors
// in reverse topological order; an A merge into a node that already holds
// one is charged gamma_1 <|A|, magnitude> (backward_bound.py:749-823), at
// once for an earlier contribution and deferred to the node's own visit for
// the last one (:801-806).  Which case applies depends on the program's
// structure only, so it is fixed here and the kernel reads it off the tape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "dag_tape.h"
#include "gearhip_internal.h"
#include "sshape.h"

namespace {

using gh::DagFwd;
using gh::DagJob;
using gh::DagPass;
using gh::DagStep;

bool is_relaxed(int op)
{
    return op == GH_OP_MUL || op == GH_OP_SQR || op == GH_OP_POW || op == GH_OP_EXP ||
           op == GH_OP_LOG || op == GH_OP_RECIP || op == GH_OP_SIN || op == GH_OP_COS;
}

bool odd_power(const gh_node &nd)
{
    return nd.op == GH_OP_POW && nd.c == std::floor(nd.c) && ((int)nd.c) % 2 == 1;
}

// Tangent tables, built once per process on the host (they depend only on
// the function): key 0 sin / cos, key p an odd power p.
const std::vector<double> &host_table(int key)
{
    static std::mutex mu;
    static std::map<int, std::vector<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double> &T = cache[key];
    if (key == 0) {
        T.resize(gh::ss::SIN_BLOCK);
        gh::ss::build_sin_tables(T.data());
    } else {
        T.resize(gh::ss::POW_BLOCK);
        gh::ss::build_pow_tables(key, T.data());
    }
    return T;
}

bool is_binary(int op) { return op == GH_OP_ADD || op == GH_OP_SUB || op == GH_OP_MUL; }

constexpr int kMaxRowNodes = 8192;

struct Builder {
    std::vector<DagJob> jobs;
    std::vector<DagFwd> fwd;
    std::vector<DagPass> passes;
    std::vector<DagStep> steps;
    std::vector<int32_t> cin;
    std::vector<int32_t> gin;
    // tangent tables the tape's sin / cos / odd-power nodes read: table key
    // -> offset (doubles) in `tabs`
    std::vector<double> tabs;
    std::map<int, int32_t> tab_off;
    int32_t table(int key)
    {
        auto it = tab_off.find(key);
        if (it != tab_off.end()) return it->second;
        const std::vector<double> &T = host_table(key);
        int32_t off = (int32_t)tabs.size();
        tabs.insert(tabs.end(), T.begin(), T.end());
        tab_off[key] = off;
        return off;
    }
    // per program: state words, I/O words and backward steps per box
    std::vector<int64_t> prog_words, prog_io, prog_steps;
    // per program, per box: algorithmic FP64 operations and transcendental
    // evaluations by the measurement's counting rule (add, sub, mul 1, fma 2;
    // min / max / compare / select / abs 0; a relaxation ~10 per element,
    // the A scaling 4 and the A67 enclosure 2 per element per side,
    // concretization 4 per input per side, interval propagation 2 per node)
    std::vector<int64_t> prog_flops, prog_trans;

    // the counting rule applied to one pass
    int64_t pass_flops(int pidx) const
    {
        const DagPass &P = passes[pidx];
        const int sides = P.want_u ? 2 : 1;
        int64_t f = 0;
        for (int k = 0; k < P.nsteps; ++k) {
            const DagStep &st = steps[P.step_off + k];
            if (st.op == GH_OP_INPUT || st.op == GH_OP_CONST) {
                f += 2 * sides;
                continue;
            }
            const int pushes = is_binary(st.op) ? 2 : 1;
            f += sides * pushes * (4 + 2);   // A scaling and enclosure per receiving input
        }
        return f + (int64_t)4 * P.ncin * sides;
    }

    // 8-byte slot accesses per lane of one pass, as the kernel performs them
    int64_t pass_words(int pidx) const
    {
        const DagPass &P = passes[pidx];
        const int sides = P.want_u ? 2 : 1;
        auto pw = [](uint8_t kind) -> int64_t {
            return kind == gh::PUSH_FIRST ? 1 : (kind == gh::PUSH_DEFER ? 2 : 4);
        };
        int64_t w = sides;  // the start node's A
        for (int k = 0; k < P.nsteps; ++k) {
            const DagStep &st = steps[P.step_off + k];
            if (st.op == GH_OP_INPUT) {
                w += 2 * sides + 2;
                continue;
            }
            w += sides;
            if (st.op == GH_OP_CONST) continue;
            if (st.flags & gh::STEP_DEFERRED) w += 2;
            switch (st.op) {
            case GH_OP_ADD:
            case GH_OP_SUB: w += 4 + sides * (pw(st.ka) + pw(st.kb)); break;
            case GH_OP_ADDC: w += 2 + sides * pw(st.ka); break;
            case GH_OP_NEG: w += sides * pw(st.ka); break;
            case GH_OP_MULC: w += 2 + sides * pw(st.ka); break;
            case GH_OP_MUL: w += 7 + sides * (pw(st.ka) + pw(st.kb)); break;
            default: w += 5 + sides * pw(st.ka); break;
            }
        }
        return w + (int64_t)P.ncin * (2 + sides);
    }

    struct Local {
        int op, a, b;
        double c;
        int gid;
    };

    // A-slot allocator of one pass (pairs: lower at s, upper at s + 1)
    struct Slots {
        int base, used = 0, peak = 0;
        std::vector<int> free_list;
        int get()
        {
            int s;
            if (!free_list.empty()) {
                s = free_list.back();
                free_list.pop_back();
            } else {
                s = used++;
            }
            peak = std::max(peak, used);
            return base + 2 * s;
        }
        void put(int slot) { free_list.push_back((slot - base) / 2); }
    };

    // one backward pass from local node s; returns the pass index
    int add_pass(const std::vector<Local> &N, int ni, int s, int want_u, int abase, int *apeak,
                 const std::vector<int> &rs)
    {
        const int V = (int)N.size();
        std::vector<char> anc(s + 1, 0), has(s + 1, 0), df(s + 1, 0);
        std::vector<int> rem(s + 1, 0), slot(s + 1, -1);
        anc[s] = 1;
        for (int v = s; v >= ni; --v) {
            if (!anc[v] || N[v].op == GH_OP_CONST) continue;
            anc[N[v].a] = 1;
            rem[N[v].a]++;
            if (is_binary(N[v].op)) {
                anc[N[v].b] = 1;
                rem[N[v].b]++;
            }
        }
        Slots sl;
        sl.base = abase;
        has[s] = 1;
        slot[s] = sl.get();
        DagPass P{};
        P.step_off = (int32_t)steps.size();
        P.start = s;
        P.want_u = want_u;
        P.sstart = slot[s];
        auto push = [&](int u, int later) -> uint8_t {
            if (has[u]) {
                bool last = !later && rem[u] == 0;
                if (last) {
                    df[u] = 1;
                    return gh::PUSH_DEFER;
                }
                return gh::PUSH_CHARGE;
            }
            has[u] = 1;
            slot[u] = sl.get();
            return gh::PUSH_FIRST;
        };
        bool first_input = true;
        std::vector<std::pair<int, int>> ins;
        for (int v = s; v >= 0; --v) {
            if (!anc[v]) continue;
            const Local &nd = N[v];
            DagStep st{};
            st.op = (uint8_t)nd.op;
            st.v = (int16_t)v;
            st.sv = (int16_t)slot[v];
            st.c = nd.c;
            st.rs = (int16_t)(rs[v] >= 0 ? rs[v] : 0);
            if (nd.op == GH_OP_INPUT) {
                if (first_input) st.flags |= gh::STEP_FIRST_INPUT;
                first_input = false;
                ins.emplace_back(v, slot[v]);
                steps.push_back(st);
                continue;  // its A stays live until concretization
            }
            if (nd.op == GH_OP_CONST) {
                steps.push_back(st);
                sl.put(slot[v]);
                continue;
            }
            if (df[v]) st.flags |= gh::STEP_DEFERRED;
            int a = nd.a, b = nd.b;
            rem[a]--;
            if (is_binary(nd.op)) rem[b]--;
            st.a = (int16_t)a;
            st.b = (int16_t)b;
            switch (nd.op) {
            case GH_OP_ADD:
            case GH_OP_SUB:
            case GH_OP_MUL:
                st.ka = push(a, b == a);
                st.kb = push(b, 0);
                st.sb = (int16_t)slot[b];
                break;
            default:
                st.ka = push(a, 0);
                break;
            }
            st.sa = (int16_t)slot[a];
            steps.push_back(st);
            sl.put(slot[v]);
        }
        std::sort(ins.begin(), ins.end());
        P.cin_off = (int32_t)cin.size();
        P.ncin = (int32_t)ins.size();
        for (auto &p : ins) cin.push_back((p.first << 16) | p.second);
        P.nsteps = (int32_t)steps.size() - P.step_off;
        *apeak = std::max(*apeak, sl.peak);
        passes.push_back(P);
        (void)V;
        return (int)passes.size() - 1;
    }

    int add_program(int p, const gh_prog &G)
    {
        prog_words.push_back(0);
        prog_io.push_back(0);
        prog_steps.push_back(0);
        prog_flops.push_back(0);
        prog_trans.push_back(0);
        const int n = G.n, Vg = G.V;
        if (n < 0 || Vg < n || G.R < 0 || (Vg > 0 && !G.nodes) || (G.R > 0 && !G.out))
            return GH_EINVAL;
        // inputs of relaxed nodes get CROWN intermediate bounds
        std::vector<char> needs(Vg, 0);
        for (int v = n; v < Vg; ++v) {
            const gh_node &nd = G.nodes[v];
            if (nd.op < GH_OP_CONST || nd.op > GH_OP_COS) return GH_EINVAL;
            if (nd.op != GH_OP_CONST && (nd.a < 0 || nd.a >= v)) return GH_EINVAL;
            if (is_binary(nd.op) && (nd.b < 0 || nd.b >= v)) return GH_EINVAL;
            if (nd.op == GH_OP_POW) {
                // integers 3 .. 31, or non-integers > 0 (x ** 2 is SQR)
                if (!(nd.c > 0) || nd.c > 31) return GH_EUNSUPPORTED;
                if (nd.c == std::floor(nd.c) && nd.c < 3) return GH_EUNSUPPORTED;
            }
            if (!is_relaxed(nd.op)) continue;
            if (nd.a >= n) needs[nd.a] = 1;
            if (nd.op == GH_OP_MUL && nd.b >= n) needs[nd.b] = 1;
        }
        for (int v = 0; v < n; ++v)
            if (G.nodes[v].op != GH_OP_INPUT || G.nodes[v].a != v) return GH_EINVAL;
        for (int r = 0; r < G.R; ++r) {
            int o = G.out[r];
            if (o < 0 || o >= Vg) return GH_EINVAL;
            // the row's ancestors
            std::vector<char> in_row(o + 1, 0);
            std::vector<int> stack{o};
            while (!stack.empty()) {
                int v = stack.back();
                stack.pop_back();
                if (in_row[v]) continue;
                in_row[v] = 1;
                const gh_node &nd = G.nodes[v];
                if (nd.op == GH_OP_INPUT || nd.op == GH_OP_CONST) continue;
                stack.push_back(nd.a);
                if (is_binary(nd.op)) stack.push_back(nd.b);
            }
            std::vector<int> order, local(o + 1, -1);
            for (int v = 0; v < std::min(n, o + 1); ++v)
                if (in_row[v]) order.push_back(v);
            int ni = (int)order.size();
            for (int v = n; v <= o; ++v)
                if (in_row[v]) order.push_back(v);
            const int V = (int)order.size();
            if (V > kMaxRowNodes) return GH_EUNSUPPORTED;
            for (int k = 0; k < V; ++k) local[order[k]] = k;
            std::vector<Local> N(V);
            for (int k = 0; k < V; ++k) {
                const gh_node &nd = G.nodes[order[k]];
                N[k].op = nd.op;
                N[k].c = nd.c;
                N[k].gid = order[k];
                N[k].a = (nd.op == GH_OP_INPUT || nd.op == GH_OP_CONST) ? 0 : local[nd.a];
                N[k].b = is_binary(nd.op) ? local[nd.b] : 0;
            }
            DagJob J{};
            J.prog = p;
            J.row = r;
            J.V = V;
            J.ni = ni;
            J.n = n;
            J.R = G.R;
            J.out = local[o];
            J.gin_off = (int32_t)gin.size();
            for (int k = 0; k < ni; ++k) gin.push_back(order[k]);
            // relaxation slots after L (0..V-1) and U (V..2V-1)
            std::vector<int> rs(V, -1);
            int next = 2 * V;
            for (int k = ni; k < V; ++k) {
                if (!is_relaxed(N[k].op)) continue;
                rs[k] = next;
                next += (N[k].op == GH_OP_MUL) ? 7 : 5;
            }
            const int abase = next;
            int apeak = 0;
            J.fwd_off = (int32_t)fwd.size();
            J.pass_off = (int32_t)passes.size();
            int64_t cost = V;
            int64_t words = 2 * (int64_t)ni;  // the row's inputs into L / U
            int64_t flops = 0, trans = 0;
            for (int k = ni; k < V; ++k) {
                DagFwd f{};
                f.op = N[k].op;
                f.a = N[k].a;
                f.b = N[k].b;
                f.rs = rs[k];
                f.c = N[k].c;
                f.pass = -1;
                if (N[k].op == GH_OP_SIN || N[k].op == GH_OP_COS) f.tab = table(0);
                else if (odd_power(G.nodes[N[k].gid])) f.tab = table((int)N[k].c);
                words += 4 + (is_binary(f.op) ? 2 : 0) + (rs[k] < 0 ? 0 : (f.op == GH_OP_MUL ? 7 : 5));
                flops += 2 + (rs[k] < 0 ? 0 : 10);
                if (f.op == GH_OP_EXP) trans += 5;        // interval 2, planes 3
                else if (f.op == GH_OP_LOG) trans += 5;
                else if (f.op == GH_OP_SIN || f.op == GH_OP_COS) trans += 8;
                else if (f.op == GH_OP_POW && f.c != std::floor(f.c)) trans += 6;
                if (needs[N[k].gid] && N[k].op != GH_OP_CONST) {
                    f.pass = add_pass(N, ni, k, 1, abase, &apeak, rs);
                    cost += passes.back().nsteps;
                    words += pass_words(f.pass) + 2;
                    flops += pass_flops(f.pass);
                }
                fwd.push_back(f);
            }
            if (o < n) {
                J.final_pass = -1;
            } else {
                J.final_pass = add_pass(N, ni, J.out, 0, abase, &apeak, rs);
                cost += passes.back().nsteps;
                words += pass_words(J.final_pass);
                flops += pass_flops(J.final_pass);
            }
            prog_words[p] += words;
            prog_flops[p] += flops;
            prog_trans[p] += trans;
            prog_io[p] += 2 * (int64_t)ni + 1;  // box endpoints in, bound out
            prog_steps[p] += cost - V;
            J.slots = abase + 2 * apeak;
            J.abase = abase;
            if (J.slots > 32767) return GH_EUNSUPPORTED;
            J.cost = cost;
            jobs.push_back(J);
        }
        return GH_OK;
    }
};

template <class T>
size_t append(std::vector<char> &buf, const std::vector<T> &v)
{
    size_t off = (buf.size() + 255) & ~size_t(255);
    buf.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

}  // namespace

extern "C" int gh_dag_plan_create(int nprog, const gh_prog *progs, gh_dag_plan **plan)
{
    if (!plan || nprog < 0 || (nprog > 0 && !progs)) return GH_EINVAL;
    *plan = nullptr;
    Builder B;
    for (int p = 0; p < nprog; ++p) {
        int rc = B.add_program(p, progs[p]);
        if (rc != GH_OK) return rc;
    }
    // largest rows first: the queue hands them out before the short ones
    std::stable_sort(B.jobs.begin(), B.jobs.end(),
                     [](const DagJob &x, const DagJob &y) { return x.cost > y.cost; });
    std::vector<char> buf;
    size_t o_jobs = append(buf, B.jobs), o_fwd = append(buf, B.fwd), o_pass = append(buf, B.passes);
    size_t o_steps = append(buf, B.steps), o_cin = append(buf, B.cin), o_gin = append(buf, B.gin);
    size_t o_tabs = append(buf, B.tabs);
    buf.resize(std::max<size_t>(buf.size(), 256));
    auto *P = new gh_dag_plan{};
    if (hipGetDevice(&P->device) != hipSuccess || hipMalloc(&P->dbuf, buf.size()) != hipSuccess) {
        delete P;
        return GH_ELAUNCH;
    }
    if (hipMemcpy(P->dbuf, buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(P->dbuf);
        delete P;
        return GH_ELAUNCH;
    }
    char *d = static_cast<char *>(P->dbuf);
    P->jobs = reinterpret_cast<const DagJob *>(d + o_jobs);
    P->fwd = reinterpret_cast<const DagFwd *>(d + o_fwd);
    P->passes = reinterpret_cast<const DagPass *>(d + o_pass);
    P->steps = reinterpret_cast<const DagStep *>(d + o_steps);
    P->cin = reinterpret_cast<const int32_t *>(d + o_cin);
    P->gin = reinterpret_cast<const int32_t *>(d + o_gin);
    P->tabs = reinterpret_cast<const double *>(d + o_tabs);
    P->njobs = (int64_t)B.jobs.size();
    P->nfwd = (int64_t)B.fwd.size();
    P->npass = (int64_t)B.passes.size();
    P->nsteps = (int64_t)B.steps.size();
    P->ncin = (int64_t)B.cin.size();
    P->ngin = (int64_t)B.gin.size();
    P->nprog = nprog;
    P->max_slots = 1;
    P->max_aslots = 0;
    for (auto &J : B.jobs) {
        P->max_slots = std::max<int64_t>(P->max_slots, J.slots);
        P->max_aslots = std::max<int64_t>(P->max_aslots, J.slots - J.abase);
    }
    P->tape_bytes = (int64_t)buf.size();
    for (int p = 0; p < nprog; ++p) {
        P->state_bytes.push_back(8 * B.prog_words[p]);
        P->io_bytes.push_back(8 * B.prog_io[p]);
        P->visits.push_back(B.prog_steps[p]);
        P->flops.push_back(B.prog_flops[p]);
        P->trans.push_back(B.prog_trans[p]);
    }
    *plan = P;
    return GH_OK;
}

extern "C" int gh_dag_plan_destroy(gh_dag_plan *plan)
{
    if (!plan) return GH_OK;
    hipError_t e = hipFree(plan->dbuf);
    delete plan;
    return e == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" int gh_dag_plan_info(const gh_dag_plan *plan, int64_t *jobs, int64_t *max_slots,
                                int64_t *tape_bytes)
{
    if (!plan) return GH_EINVAL;
    if (jobs) *jobs = plan->njobs;
    if (max_slots) *max_slots = plan->max_slots;
    if (tape_bytes) *tape_bytes = plan->tape_bytes;
    return GH_OK;
}

extern "C" int gh_dag_plan_stats(const gh_dag_plan *plan, int64_t *state_bytes, int64_t *io_bytes,
                                 int64_t *steps)
{
    if (!plan) return GH_EINVAL;
    for (int64_t p = 0; p < plan->nprog; ++p) {
        if (state_bytes) state_bytes[p] = plan->state_bytes[p];
        if (io_bytes) io_bytes[p] = plan->io_bytes[p];
        if (steps) steps[p] = plan->visits[p];
    }
    return GH_OK;
}

extern "C" int gh_dag_plan_work(const gh_dag_plan *plan, int64_t *flops, int64_t *transcendentals)
{
    if (!plan) return GH_EINVAL;
    for (int64_t p = 0; p < plan->nprog; ++p) {
        if (flops) flops[p] = plan->flops[p];
        if (transcendentals) transcendentals[p] = plan->trans[p];
    }
    return GH_OK;
}

extern "C" int gh_dag_plan_aslots(const gh_dag_plan *plan, int64_t *max_aslots)
{
    if (!plan || !max_aslots) return GH_EINVAL;
    *max_aslots = plan->max_aslots;
    return GH_OK;
}
