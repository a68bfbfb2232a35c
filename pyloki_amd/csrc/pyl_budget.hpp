// Internal (not C-ABI) entry points shared by the EP level loops, the survivor budget and
// the ranking network's translation units.
#pragma once

#include "pyl_common.hpp"

namespace pyl {

// The budget of search level `level`: the schedule's entry, or the single budget.
inline int64_t budget_of(const pyl_ep_budget* b, int level) {
  return b->level_budget != nullptr ? b->level_budget[level] : b->budget;
}
// The ranking key of search level `level` (the level's entry of level_kind, else kind).
inline int kind_of(const pyl_ep_budget* b, int level) {
  return b->level_kind != nullptr ? b->level_kind[level] : b->kind;
}

// The network's forward pass over at most n_cap samples, the count read from n_dev
// when not null; bf16 selects the MFMA inference kernel.
int cnn_forward_count(const float* params, const float* folds, int64_t n_cap,
                      const int64_t* n_dev, int nbins, int complex_folds, const float* feats,
                      float* logits, int bf16, hipStream_t s);

// The survivor budget of one level. The level's record `rec` (PYL_EP_REC_WORDS words; the
// next record follows) gives the survivors (NSURV) and the first history row (HIST); the
// kept rows of leaves/folds/scores go to o_*, their backtracks to rows HIST.. of o_hist,
// and NKEPT, the next record's NPAR and HIST are written. n_cap bounds the survivors.
int ep_budget_enqueue(const pyl_ep_budget* b, int P, int level, const double* leaves,
                      const float* folds, const float* scores, const int32_t* bt,
                      int64_t n_cap, int64_t* rec, double* o_leaves, float* o_folds,
                      float* o_scores, int32_t* o_hist, hipStream_t s);

// The budget of level `level` in the level loop: ep_budget_enqueue, or, when b->level_mask
// flags the level off, one launch that copies the survivors to o_* (and their backtracks to
// the history) when they fit the budget, else stops the level with PYL_EP_ST_BUDGET.
int ep_budget_level_enqueue(const pyl_ep_budget* b, int P, int level, const double* leaves,
                            const float* folds, const float* scores, const int32_t* bt,
                            int64_t n_cap, int64_t* rec, double* o_leaves, float* o_folds,
                            float* o_scores, int32_t* o_hist, hipStream_t s);

// The budget of a forest's level (pyl_ep_forest): every tree ranked and cut on its own.
// tab is the level's tree table (its SO row read, KO written), tab_next the next level's
// (PO written); the record as in ep_budget_enqueue, with the forest's totals. With
// b->level_mask flagging the level off, one pass as in ep_budget_level_enqueue (any tree
// above the budget stops the level).
int ep_budget_forest_enqueue(const pyl_ep_budget* b, const pyl_ep_forest* F, int P, int level,
                             int64_t* tab, int64_t* tab_next, const double* leaves,
                             const float* folds, const float* scores, const int32_t* bt,
                             int64_t n_cap, int64_t* rec, double* o_leaves, float* o_folds,
                             float* o_scores, int32_t* o_hist, hipStream_t s);

// The S/N cut of a forest's level before anything is materialised: keys are the
// survivors' scores in survivor order; b->kept receives every tree's kept survivors
// (tree t's at b->kept[SO[t]..], survivor indices), tab's KO row, tab_next's PO row and the
// record's NKEPT, next NPAR and next HIST are written. With b->level_mask flagging the
// level off: every tree keeps its survivors, or the level stops with PYL_EP_ST_BUDGET.
int ep_budget_forest_select(const pyl_ep_budget* b, const pyl_ep_forest* F, int level,
                            int64_t* tab, int64_t* tab_next, const float* keys, int64_t n_cap,
                            int64_t* rec, hipStream_t s);

}  // namespace pyl
