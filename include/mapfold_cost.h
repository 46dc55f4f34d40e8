/*
 * mapfold_cost.h -- learned subtree-cost model of libmapfold.so (C ABI).
 *
 * A small perceptron (MF_COST_FEATURES -> 128 -> 128 -> 128 -> 1, ReLU) that
 * predicts log2(1 + nodes) of a work item's DFS subtree from its state.  It
 * orders and splits work only; no count ever depends on it.  Trained in FP32
 * on the GPU by the library's own forward/backward kernels (Adam), run for
 * prediction by a BF16 MFMA kernel (BF16 weights and activations, FP32
 * accumulation).
 *
 * Reference interfaces replaced (johnowhitaker/map_folding, file:line):
 *   mf_cost_features_sym + mf_cost_predict + learned order
 *       sym_fold.cpp:235-250 (estimate_work: lookahead leaf count as the cost
 *       proxy) and :267-314 (score_and_sort_cases, MF_COST_DEPTH /
 *       MF_SORT_CASES).  Here the order is a permutation handed to the solver
 *       (mf_solve_options.order), so case indices and per-case results keep
 *       the stream order.
 *   training pairs (state, node count)
 *       experiments/9x9/sym_fold9.cpp:785-848 (MF_PROFILE_WORK per-case node
 *       counts); here the GPU solve's per-item node counts (mf_frontier_solve
 *       per_nodes).
 *
 * Calls share the library's stream and context (mf_init); destroy models
 * before mf_shutdown.
 */
#ifndef MAPFOLD_COST_H
#define MAPFOLD_COST_H

#include <stdint.h>

#include "mapfold.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MF_COST_FEATURES = 80, /* 64 state encoding + 16 scalar features */
    MF_COST_HIDDEN = 128,
    MF_COST_PARAMS = 43521 /* W1[128x80] b1 W2[128x128] b2 W3[128x128] b3 w4[128] b4 */
};

typedef struct mf_cost_model mf_cost_model;

typedef struct {
    int32_t epochs;   /* passes over the data (default 20) */
    int32_t batch;    /* rows per Adam step (default 512) */
    float lr;         /* default 1e-3 */
    float beta1;      /* default 0.9 */
    float beta2;      /* default 0.999 */
    float eps;        /* default 1e-8 */
    uint64_t seed;    /* shuffle order */
} mf_cost_train_options;

/* He-uniform weights from `seed`, zero biases. */
int mf_cost_create(uint64_t seed, mf_cost_model **out);
void mf_cost_destroy(mf_cost_model *m);
int mf_cost_get_params(const mf_cost_model *m, float *params);   /* MF_COST_PARAMS floats */
int mf_cost_set_params(mf_cost_model *m, const float *params);

/* Feature rows (count x MF_COST_FEATURES floats) of sym cases at `level`
 * (64 cycle bytes per case) and of Lunnon states at partition depth `depth`
 * (depth+1 bytes b[0..depth] per state, mf_lunnon_frontier layout). */
int mf_cost_features_sym(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, float *out);
int mf_cost_features_lunnon(const int32_t *dims, int32_t ndim, int32_t depth, const uint8_t *states,
                            int64_t count, float *out);

/* One FP32 forward + backward over all rows: mean-squared-error loss and
 * its gradient (MF_COST_PARAMS floats, parameter layout). */
int mf_cost_gradient(mf_cost_model *m, const float *x, const float *y, int64_t count, float *grad, float *loss);

/* Adam training on the GPU; epoch_loss (nullable) receives the mean batch
 * loss of each epoch. */
int mf_cost_train(mf_cost_model *m, const float *x, const float *y, int64_t count,
                  const mf_cost_train_options *opt, float *epoch_loss);

/* The reference's lookahead cost proxy (estimate_work, sym_fold.cpp:235-250;
 * sym_fold9.cpp:359-374 for n = 9): for each case (case rows as in
 * mapfold.h: 64 bytes, 128 for n = 9), the number of its descendants at
 * level min(n*n, level + depth) (1 when that is the case's own level). */
int mf_sym_estimate_work(int32_t n, int32_t level, const uint8_t *cycles, int64_t count, int32_t depth,
                         uint64_t *scores);

/* Predictions: BF16 MFMA inference, and the FP32 training forward. */
int mf_cost_predict(mf_cost_model *m, const float *x, int64_t count, float *pred);
int mf_cost_predict_fp32(mf_cost_model *m, const float *x, int64_t count, float *pred);
/* The same prediction for every item of a resident sym frontier (n <= 8)
 * without a host round trip: the feature rows (the values of
 * mf_cost_features_sym) are computed on the device from the resident items,
 * then the BF16 inference runs on them.  *rows = the frontier's items;
 * pred needs room for them (max_rows). */
int mf_cost_predict_frontier(mf_cost_model *m, mf_frontier *f, float *pred, int64_t max_rows, int64_t *rows);
/* The device-computed feature rows themselves (80 floats per item). */
int mf_cost_features_frontier(mf_frontier *f, float *x, int64_t max_rows, int64_t *rows);

#ifdef __cplusplus
}
#endif

#endif /* MAPFOLD_COST_H */
