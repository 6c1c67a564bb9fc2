// Tape format of the expression-program bound kernel.
//
// A plan holds one job per (program, row).  A job is the row's local
// subprogram -- the inputs it reads (ascending global index) followed by its
// operation nodes in program order -- compiled into
//   - a forward tape: per operation node, interval bound, relaxation planes
//     and (for an input of a relaxed node) the index of its CROWN pass;
//   - backward passes: per pass, the ancestors of its start node in
//     decreasing node order, each with every bookkeeping decision of the
//     backward sweep resolved at plan time (which A values are first
//     assignments, which merges are charged at once and which are deferred
//     to the receiving node's own visit), and the pass's input nodes in
//     ascending order for concretization.
// The decisions depend on the program's structure only, so every lane of a
// wavefront (one box each) walks the same tape with wave-uniform control
// flow.  Per-lane state lives in a per-wavefront workspace laid out
// [slot][lane], so each slot access is one coalesced 512-byte line:
//   slot k          L of local node k
//   slot V + k      U of local node k
//   rs .. rs+6      relaxation planes of a relaxed node (MUL: 6 + flags,
//                   unary: 4 + flags at rs + 4)
//   sa, sa + 1      A (lower, upper) of a node during a pass: the A area,
//                   slots abase .. slots - 1, the most accessed per slot;
//                   the kernel keeps its first lds_slots slots in LDS
#ifndef GEARHIP_DAG_TAPE_H
#define GEARHIP_DAG_TAPE_H

#include <stdint.h>

#include <vector>

namespace gh {

// push kinds: how an A contribution lands in its receiving node
enum : uint8_t {
    PUSH_FIRST = 0,   // first contribution: assignment
    PUSH_DEFER = 1,   // last contribution: merge charged at the node's visit
    PUSH_CHARGE = 2,  // earlier contribution: merge charged now
};

// step flags
enum : uint8_t {
    STEP_DEFERRED = 1,  // this node's own deferred merge charge applies
    STEP_FIRST_INPUT = 2,  // first input node of the pass (starts the root sum)
};

struct DagStep {      // 32 bytes, one per visited ancestor
    uint8_t op, flags, ka, kb;
    int16_t v, a, b;  // local node ids (L slot = id, U slot = V + id)
    int16_t sv, sa, sb;  // A slots of v and of its operands
    int16_t rs, pad;  // relaxation slot of v (relaxed ops)
    int32_t pad2;
    double c;
};

struct DagFwd {       // 32 bytes, one per operation node
    int32_t op, a, b, rs;
    int32_t pass;     // CROWN pass of this node, -1 for none
    int32_t tab;      // sin / cos / odd power: offset of its tangent tables
    double c;
};

struct DagPass {
    int32_t step_off, nsteps;
    int32_t start, want_u;
    int32_t cin_off, ncin;   // concretization list: (node, A slot) pairs
    int32_t sstart, pad;     // A slot of the start node
};

struct DagJob {
    int32_t prog, row;
    int32_t V, ni;           // local nodes, local inputs
    int32_t fwd_off, pass_off;  // forward tape, first pass of the job
    int32_t final_pass;      // -1 when the row is a bare input
    int32_t out;             // local id of the row's node
    int32_t n, R;            // program inputs (box stride) and rows
    int32_t gin_off;         // global indices of the local inputs
    int32_t slots;           // per-lane slots this job uses
    int32_t abase, pad;      // first slot of the A area (the passes' A pairs)
    int64_t cost;            // backward steps + nodes (scheduling weight)
};

}  // namespace gh

// Host handle of a plan: device copies of the tape sections.
struct gh_dag_plan {
    int device;
    int64_t njobs, nfwd, npass, nsteps, ncin, ngin, nprog;
    int64_t max_slots, tape_bytes;
    int64_t max_aslots;         // largest A area of a job (slots)
    void *dbuf;                 // one device allocation holding every section
    const gh::DagJob *jobs;
    const gh::DagFwd *fwd;
    const gh::DagPass *passes;
    const gh::DagStep *steps;
    const int32_t *cin;         // (node << 16) | A slot
    const int32_t *gin;
    const double *tabs;         // tangent tables (gear_amd/csrc/sshape.h)
    // per program, per box: bytes of per-lane state the tape reads and
    // writes (the A propagation and relaxation data), and box / bound I/O
    std::vector<int64_t> state_bytes, io_bytes, visits;
    // per program, per box: algorithmic FP64 operations and transcendentals
    std::vector<int64_t> flops, trans;
};

#endif
