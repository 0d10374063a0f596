/*
 * opcount_rt.c -- TEST INFRASTRUCTURE: run-time side of the op-counting
 * build of oracle/mop_oracle.c (oracle/opcount.py inserts the calls).
 * Counts how often each basic block runs, per phase; the static table of
 * floating-point operations per block turns that into operations per frame.
 * Single-threaded, like the restatement.
 */
#include <stdint.h>
#include <string.h>

#define OPC_PHASES 11 /* len(opcount.PHASES) */

extern uint64_t mop_opc_counts[];
extern const int mop_opc_nblocks;
static int g_phase;

void mop_opc_bb(int id) { mop_opc_counts[(size_t)g_phase * (size_t)mop_opc_nblocks + id]++; }

int mop_opc_enter(int phase) {
  int old = g_phase;
  g_phase = phase;
  return old;
}

void mop_opc_leave(int phase) { g_phase = phase; }

int mop_opcount_blocks(void) { return mop_opc_nblocks; }
int mop_opcount_phases(void) { return OPC_PHASES; }

/* phases x blocks run counts, phase-major */
void mop_opcount_read(uint64_t *out) {
  memcpy(out, mop_opc_counts, sizeof(uint64_t) * OPC_PHASES * (size_t)mop_opc_nblocks);
}

void mop_opcount_reset(void) {
  memset(mop_opc_counts, 0, sizeof(uint64_t) * OPC_PHASES * (size_t)mop_opc_nblocks);
}
