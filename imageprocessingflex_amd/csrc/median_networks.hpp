// GENERATED by gen_median_networks.py -- do not edit.
// Batcher odd-even merge networks, padded with +inf wires, lowered and
// pruned to the wires that reach the output; verified exhaustively by
// the 0-1 principle in the generator.
// sort5: 9 comparators; med25_cols: 82 comparators, 140 min/max ops.
#pragma once

template <class K>
__device__ __forceinline__ void sort5(K (&v)[5]) {
    { const K a = v[0], b = v[1]; v[0] = kmin(a, b); v[1] = kmax(a, b); }
    { const K a = v[2], b = v[3]; v[2] = kmin(a, b); v[3] = kmax(a, b); }
    { const K a = v[0], b = v[2]; v[0] = kmin(a, b); v[2] = kmax(a, b); }
    { const K a = v[1], b = v[3]; v[1] = kmin(a, b); v[3] = kmax(a, b); }
    { const K a = v[1], b = v[2]; v[1] = kmin(a, b); v[2] = kmax(a, b); }
    { const K a = v[0], b = v[4]; v[0] = kmin(a, b); v[4] = kmax(a, b); }
    { const K a = v[2], b = v[4]; v[2] = kmin(a, b); v[4] = kmax(a, b); }
    { const K a = v[1], b = v[2]; v[1] = kmin(a, b); v[2] = kmax(a, b); }
    { const K a = v[3], b = v[4]; v[3] = kmin(a, b); v[4] = kmax(a, b); }
}

template <class K>
__device__ __forceinline__ K med25_cols(const K (&v)[25]) {
    K w0 = v[0];
    K w1 = v[1];
    K w2 = v[2];
    K w3 = v[3];
    K w4 = v[4];
    K w5;
    K w6;
    K w7;
    K w8 = v[5];
    K w9 = v[6];
    K w10 = v[7];
    K w11 = v[8];
    K w12 = v[9];
    K w13;
    K w14;
    K w15;
    K w16 = v[10];
    K w17 = v[11];
    K w18 = v[12];
    K w19 = v[13];
    K w20 = v[14];
    K w21;
    K w22;
    K w23;
    K w24 = v[15];
    K w25 = v[16];
    K w26 = v[17];
    K w27 = v[18];
    K w28 = v[19];
    K w32 = v[20];
    K w33 = v[21];
    K w34 = v[22];
    K w35 = v[23];
    K w36 = v[24];
    { const K a = w0, b = w8; w0 = kmin(a, b); w8 = kmax(a, b); }
    { const K a = w4, b = w12; w4 = kmin(a, b); w12 = kmax(a, b); }
    { const K a = w4, b = w8; w4 = kmin(a, b); w8 = kmax(a, b); }
    { const K a = w2, b = w10; w2 = kmin(a, b); w10 = kmax(a, b); }
    w6 = w10;
    { const K a = w2, b = w4; w2 = kmin(a, b); w4 = kmax(a, b); }
    { const K a = w6, b = w8; w6 = kmin(a, b); w8 = kmax(a, b); }
    w10 = w12;
    { const K a = w1, b = w9; w1 = kmin(a, b); w9 = kmax(a, b); }
    w5 = w9;
    { const K a = w3, b = w11; w3 = kmin(a, b); w11 = kmax(a, b); }
    w7 = w11;
    { const K a = w3, b = w5; w3 = kmin(a, b); w5 = kmax(a, b); }
    { const K a = w1, b = w2; w1 = kmin(a, b); w2 = kmax(a, b); }
    { const K a = w3, b = w4; w3 = kmin(a, b); w4 = kmax(a, b); }
    { const K a = w5, b = w6; w5 = kmin(a, b); w6 = kmax(a, b); }
    { const K a = w7, b = w8; w7 = kmin(a, b); w8 = kmax(a, b); }
    w9 = w10;
    { const K a = w16, b = w24; w16 = kmin(a, b); w24 = kmax(a, b); }
    { const K a = w20, b = w28; w20 = kmin(a, b); w28 = kmax(a, b); }
    { const K a = w20, b = w24; w20 = kmin(a, b); w24 = kmax(a, b); }
    { const K a = w18, b = w26; w18 = kmin(a, b); w26 = kmax(a, b); }
    w22 = w26;
    { const K a = w18, b = w20; w18 = kmin(a, b); w20 = kmax(a, b); }
    { const K a = w22, b = w24; w22 = kmin(a, b); w24 = kmax(a, b); }
    w26 = w28;
    { const K a = w17, b = w25; w17 = kmin(a, b); w25 = kmax(a, b); }
    w21 = w25;
    { const K a = w19, b = w27; w19 = kmin(a, b); w27 = kmax(a, b); }
    w23 = w27;
    { const K a = w19, b = w21; w19 = kmin(a, b); w21 = kmax(a, b); }
    { const K a = w17, b = w18; w17 = kmin(a, b); w18 = kmax(a, b); }
    { const K a = w19, b = w20; w19 = kmin(a, b); w20 = kmax(a, b); }
    { const K a = w21, b = w22; w21 = kmin(a, b); w22 = kmax(a, b); }
    { const K a = w23, b = w24; w23 = kmin(a, b); w24 = kmax(a, b); }
    w25 = w26;
    { const K a = w0, b = w16; w0 = kmin(a, b); w16 = kmax(a, b); }
    { const K a = w8, b = w24; w8 = kmin(a, b); w24 = kmax(a, b); }
    { const K a = w8, b = w16; w8 = kmin(a, b); w16 = kmax(a, b); }
    { const K a = w4, b = w20; w4 = kmin(a, b); w20 = kmax(a, b); }
    w12 = w20;
    { const K a = w4, b = w8; w4 = kmin(a, b); w8 = kmax(a, b); }
    { const K a = w12, b = w16; w12 = kmin(a, b); w16 = kmax(a, b); }
    w20 = w24;
    { const K a = w2, b = w18; w2 = kmin(a, b); w18 = kmax(a, b); }
    w10 = w18;
    { const K a = w6, b = w22; w6 = kmin(a, b); w22 = kmax(a, b); }
    w14 = w22;
    { const K a = w6, b = w10; w6 = kmin(a, b); w10 = kmax(a, b); }
    { const K a = w2, b = w4; w2 = kmin(a, b); w4 = kmax(a, b); }
    { const K a = w6, b = w8; w6 = kmin(a, b); w8 = kmax(a, b); }
    { const K a = w10, b = w12; w10 = kmin(a, b); w12 = kmax(a, b); }
    { const K a = w14, b = w16; w14 = kmin(a, b); w16 = kmax(a, b); }
    w18 = w20;
    { const K a = w1, b = w17; w1 = kmin(a, b); w17 = kmax(a, b); }
    { const K a = w9, b = w25; w9 = kmin(a, b); w25 = kmax(a, b); }
    { const K a = w9, b = w17; w9 = kmin(a, b); w17 = kmax(a, b); }
    { const K a = w5, b = w21; w5 = kmin(a, b); w21 = kmax(a, b); }
    w13 = w21;
    { const K a = w5, b = w9; w5 = kmin(a, b); w9 = kmax(a, b); }
    { const K a = w13, b = w17; w13 = kmin(a, b); w17 = kmax(a, b); }
    w21 = w25;
    { const K a = w3, b = w19; w3 = kmin(a, b); w19 = kmax(a, b); }
    w11 = w19;
    { const K a = w7, b = w23; w7 = kmin(a, b); w23 = kmax(a, b); }
    w15 = w23;
    { const K a = w7, b = w11; w7 = kmin(a, b); w11 = kmax(a, b); }
    { const K a = w3, b = w5; w3 = kmin(a, b); w5 = kmax(a, b); }
    { const K a = w7, b = w9; w7 = kmin(a, b); w9 = kmax(a, b); }
    { const K a = w11, b = w13; w11 = kmin(a, b); w13 = kmax(a, b); }
    { const K a = w15, b = w17; w15 = kmin(a, b); w17 = kmax(a, b); }
    w19 = w21;
    { const K a = w1, b = w2; w1 = kmin(a, b); w2 = kmax(a, b); }
    { const K a = w3, b = w4; w3 = kmin(a, b); w4 = kmax(a, b); }
    { const K a = w5, b = w6; w6 = kmax(a, b); }
    { const K a = w7, b = w8; w7 = kmin(a, b); w8 = kmax(a, b); }
    { const K a = w9, b = w10; w9 = kmin(a, b); w10 = kmax(a, b); }
    { const K a = w11, b = w12; w11 = kmin(a, b); w12 = kmax(a, b); }
    { const K a = w13, b = w14; w13 = kmin(a, b); }
    { const K a = w15, b = w16; w16 = kmax(a, b); }
    { const K a = w17, b = w18; w17 = kmin(a, b); w18 = kmax(a, b); }
    { const K a = w0, b = w32; w32 = kmax(a, b); }
    { const K a = w16, b = w32; w16 = kmin(a, b); }
    { const K a = w8, b = w16; w16 = kmax(a, b); }
    { const K a = w4, b = w36; w36 = kmax(a, b); }
    w20 = w36;
    { const K a = w12, b = w20; w12 = kmin(a, b); }
    { const K a = w12, b = w16; w12 = kmin(a, b); }
    { const K a = w2, b = w34; w34 = kmax(a, b); }
    { const K a = w18, b = w34; w18 = kmin(a, b); }
    { const K a = w10, b = w18; w10 = kmin(a, b); }
    { const K a = w6, b = w10; w10 = kmax(a, b); }
    { const K a = w10, b = w12; w12 = kmax(a, b); }
    { const K a = w1, b = w33; w33 = kmax(a, b); }
    { const K a = w17, b = w33; w17 = kmin(a, b); }
    { const K a = w9, b = w17; w17 = kmax(a, b); }
    { const K a = w13, b = w17; w13 = kmin(a, b); }
    { const K a = w3, b = w35; w35 = kmax(a, b); }
    { const K a = w19, b = w35; w19 = kmin(a, b); }
    { const K a = w11, b = w19; w11 = kmin(a, b); }
    { const K a = w7, b = w11; w11 = kmax(a, b); }
    { const K a = w11, b = w13; w11 = kmin(a, b); }
    { const K a = w11, b = w12; w12 = kmax(a, b); }
    return w12;
}
