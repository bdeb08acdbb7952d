// Learned preconditioner factor: a message-passing network on the pruned
// lubrication graph that predicts the block factor used in place of IC(0).
//
// The factor has the block LDL^T form of IC(0) on IC(0)'s own pattern (the
// pruned, RCM-permuted particle graph): (D + W) D^-1 (D + W)^T ~ R_FU^nf + lambda S,
// with W on the lower blocks of the pattern and D one symmetric 6x6 pivot per
// particle; its lower-triangular factor is L_PP = chol(D_P), L_PQ = W_PQ L_QQ^-T.
// The network predicts, per lower pair, multipliers of the pair's five
// resistance coefficients in W (an equivariant basis: r r, I - r r, eps.r) and
// the weight of the pair's fill term in D, and per particle an isotropic pivot
// shift.  With zero outputs the factor is block symmetric Gauss-Seidel.
//
// Network: 3 message-passing layers of width 64 (edge message from both end
// particles and the edge state, residual edge and particle states, messages
// summed at both end particles), BF16 weights and activations into FP32-
// accumulating MFMA products (v_mfma_f32_16x16x32_bf16), SiLU activations.
#pragma once
#include <stdint.h>

namespace lp {
constexpr int H = 64;        // hidden width
constexpr int FIN = 32;      // input features per edge / particle (padded)
constexpr int OE = 32;       // edge head outputs (7 used)
constexpr int ON = 32;       // particle head outputs (2 used)
constexpr int NLAYER = 10;
constexpr int NMP = 3;       // message-passing layers
constexpr int NPASS = 4;     // parallel passes of the pivot recurrence
enum { L_ENC_N = 0, L_ENC_E = 1, L_MSG = 2, /* L_MSG + 2 l, L_UPD + 2 l */ L_UPD = 3, L_HEAD_E = 8, L_HEAD_N = 9 };
struct Layer {
    int ki, no, act;         // inputs, outputs, SiLU (1) or linear (0)
    int64_t w_off, b_off;    // offsets into the FP32 parameter vector (W row-major [ki][no], then b)
    int64_t bf_off;          // offset of the layer's BF16 copies (transposed [no][ki] and natural [ki][no])
};
void layers(Layer out[NLAYER], int64_t *nparam, int64_t *nbf);
}  // namespace lp

struct LearnedState {
    lp::Layer L[lp::NLAYER];
    int64_t nparam = 0, nbf = 0;
    float *d_w = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;   // FP32 master, gradient, Adam
    float *d_wbak = nullptr;                     // parameters and moments before the last Adam step
    uint16_t *d_wt = nullptr, *d_wn = nullptr;   // BF16 copies: [no][ki] (forward), [ki][no] (input gradients)
    bool have_w = false;
    int64_t ecap = 0;                            // edges (lower blocks) the edge buffers hold
    int ncap = 0;
    // particle buffers
    uint16_t *xn0 = nullptr, *xu[lp::NMP] = {}, *xh = nullptr;
    float *zn0 = nullptr, *h[lp::NMP + 1] = {}, *zu[lp::NMP] = {}, *agg = nullptr, *outn = nullptr;
    float *Dinv = nullptr, *Dtinv = nullptr, *D = nullptr;
    // edge buffers
    uint16_t *xe0 = nullptr, *xm[lp::NMP + 1] = {};
    float *ze0 = nullptr, *e[lp::NMP + 1] = {}, *zm[lp::NMP] = {}, *m[lp::NMP] = {}, *oute = nullptr;
    float *geo = nullptr, *W = nullptr;
    int *erow = nullptr;                         // row particle of every lower block
    int *d_fail = nullptr, *h_fail = nullptr;   // [0] failure bits, [1] pivots above level 0, [2] at level 4
    int *lvlp = nullptr;                         // [NPASS][N]: the level of each pivot's safeguard in each pass
    float *Dpinv = nullptr;                      // [NPASS - 1][N][36]: inverse pivots of the passes before the last
    float *Dbar2 = nullptr;                      // a pass's pivot gradient (backward scratch)
    float lambda = 1.0f;                         // pivot shift of the last factor (doubled until A_PP + lambda S > 0)
    int shrunk = 0, plain = 0;                   // pivots of the last factor above level 0 / at level 4
    // backward buffers
    float *gh = nullptr, *gh2 = nullptr, *ge = nullptr, *gm = nullptr, *gzn = nullptr, *gze = nullptr;
    float *gxe = nullptr, *gxn = nullptr, *goute = nullptr, *goutn = nullptr, *gW = nullptr, *gDinv = nullptr;
    uint16_t *gbf = nullptr;                     // BF16 copy of a pre-activation gradient
    float *part = nullptr; int64_t part_cap = 0; // per-chunk partial weight gradients
    // Jacobi iterates of the training loss (allocated for the sample count in use)
    float *jv = nullptr; int64_t jv_cap = 0;
    double *d_loss = nullptr, *h_loss = nullptr;
    bool active = false;                         // this step's preconditioner is the learned factor
};
