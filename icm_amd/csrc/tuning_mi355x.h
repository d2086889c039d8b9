// tuning_mi355x.h -- planner table for MI355X (gfx950), written by
// icm_amd/autotune.py from timings on the device.  First row whose limits
// cover (n, k) wins.
#pragma once

struct TuneRow {
    int n_max, k_max;
    int B;          // block size (players per leaf block)
    int sb_max;     // schoolbook for levels with c_{l-1} <= sb_max
};

static const TuneRow kTuneTable[] = {
    {2560, 5, 32, 33},   // tuned at n=1024 k=1: 0.171 ms
    {2560, 55, 32, 257},   // tuned at n=1024 k=9: 0.185 ms
    {2560, 563, 32, 129},   // tuned at n=1024 k=102: 0.237 ms
    {2560, 2147483647, 64, 512},   // tuned at n=1024 k=1024: 0.297 ms
    {10240, 5, 32, 65},   // tuned at n=4096 k=1: 0.184 ms
    {10240, 209, 64, 512},   // tuned at n=4096 k=9: 0.212 ms
    {10240, 2252, 64, 65},   // tuned at n=4096 k=409: 0.539 ms
    {10240, 2147483647, 64, 129},   // tuned at n=4096 k=4096: 0.552 ms
    {40960, 5, 64, 129},   // tuned at n=16384 k=1: 0.218 ms
    {40960, 823, 64, 129},   // tuned at n=16384 k=9: 0.257 ms
    {40960, 9011, 64, 33},   // tuned at n=16384 k=1638: 0.960 ms
    {40960, 2147483647, 64, 33},   // tuned at n=16384 k=16384: 1.204 ms
    {163840, 5, 32, 512},   // tuned at n=65536 k=1: 0.357 ms
    {163840, 3281, 64, 512},   // tuned at n=65536 k=9: 0.481 ms
    {163840, 36044, 64, 129},   // tuned at n=65536 k=6553: 4.821 ms
    {163840, 2147483647, 64, 65},   // tuned at n=65536 k=65536: 5.643 ms
    {655360, 5, 32, 65},   // tuned at n=262144 k=1: 1.187 ms
    {655360, 13111, 64, 512},   // tuned at n=262144 k=9: 1.606 ms
    {655360, 144179, 64, 129},   // tuned at n=262144 k=26214: 23.431 ms
    {655360, 2147483647, 64, 129},   // tuned at n=262144 k=262144: 26.560 ms
    {2147483647, 5, 32, 129},   // tuned at n=1048576 k=1: 4.514 ms
    {2147483647, 52433, 64, 257},   // tuned at n=1048576 k=9: 6.148 ms
    {2147483647, 576716, 64, 129},   // tuned at n=1048576 k=104857: 110.938 ms
    {2147483647, 2147483647, 64, 129},   // tuned at n=1048576 k=1048576: 123.523 ms
};

// Ragged batches (one setting per batch), tuned on the 4096-instance ragged mix.
struct BatchTune {
    int B, sb_max;
};
static const BatchTune kBatchTune = {64, 129};   // 44.259 ms
