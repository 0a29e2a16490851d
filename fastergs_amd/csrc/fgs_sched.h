// Tile cost model: an MLP 8 -> 64 -> 64 -> 2 with ReLU hidden layers.
// Parameter layout (fp32, contiguous): W1 [64][8], b1 [64], W2 [64][64],
// b2 [64], W3 [2][64], b3 [2]; W[j][k] maps input k to output j.
#pragma once

namespace fgs {

constexpr int MLP_IN = 8;
constexpr int MLP_HID = 64;
constexpr int MLP_OUT = 2;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_HID * MLP_IN;
constexpr int MLP_W2 = MLP_B1 + MLP_HID;
constexpr int MLP_B2 = MLP_W2 + MLP_HID * MLP_HID;
constexpr int MLP_W3 = MLP_B2 + MLP_HID;
constexpr int MLP_B3 = MLP_W3 + MLP_OUT * MLP_HID;
constexpr int MLP_NPARAMS = MLP_B3 + MLP_OUT;  // 4866

// Per training row the gradient pass keeps [h1 | h2 | gz1 | gz2 | gy]
// (post-ReLU activations, pre-activation gradients, output gradients).
constexpr int MLP_ACT_STRIDE = 4 * MLP_HID + MLP_OUT;
// Offsets in the concatenated row view [x | h1 | h2 | gz1 | gz2 | gy].
constexpr int MLP_XOFF_X = 0;
constexpr int MLP_XOFF_H1 = MLP_IN;
constexpr int MLP_XOFF_H2 = MLP_XOFF_H1 + MLP_HID;
constexpr int MLP_XOFF_GZ1 = MLP_XOFF_H2 + MLP_HID;
constexpr int MLP_XOFF_GZ2 = MLP_XOFF_GZ1 + MLP_HID;
constexpr int MLP_XOFF_GY = MLP_XOFF_GZ2 + MLP_HID;

}  // namespace fgs
