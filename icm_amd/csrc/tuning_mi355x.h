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
    {2560, 5, 32, 33},   // tuned at n=1024 k=1: 0.175 ms
    {2560, 55, 32, 33},   // tuned at n=1024 k=9: 0.182 ms
    {2560, 563, 32, 129},   // tuned at n=1024 k=102: 0.220 ms
    {2560, 2147483647, 64, 257},   // tuned at n=1024 k=1024: 0.263 ms
    {10240, 5, 32, 512},   // tuned at n=4096 k=1: 0.173 ms
    {10240, 209, 64, 512},   // tuned at n=4096 k=9: 0.208 ms
    {10240, 2252, 64, 33},   // tuned at n=4096 k=409: 0.475 ms
    {10240, 2147483647, 64, 65},   // tuned at n=4096 k=4096: 0.497 ms
    {40960, 5, 64, 512},   // tuned at n=16384 k=1: 0.213 ms
    {40960, 823, 64, 129},   // tuned at n=16384 k=9: 0.255 ms
    {40960, 9011, 64, 33},   // tuned at n=16384 k=1638: 0.782 ms
    {40960, 2147483647, 64, 33},   // tuned at n=16384 k=16384: 1.033 ms
    {163840, 5, 32, 129},   // tuned at n=65536 k=1: 0.356 ms
    {163840, 3281, 64, 129},   // tuned at n=65536 k=9: 0.481 ms
    {163840, 36044, 64, 17},   // tuned at n=65536 k=6553: 4.176 ms
    {163840, 2147483647, 64, 33},   // tuned at n=65536 k=65536: 4.986 ms
    {655360, 5, 32, 129},   // tuned at n=262144 k=1: 1.188 ms
    {655360, 13111, 64, 257},   // tuned at n=262144 k=9: 1.597 ms
    {655360, 144179, 64, 33},   // tuned at n=262144 k=26214: 20.688 ms
    {655360, 2147483647, 64, 0},   // tuned at n=262144 k=262144: 23.655 ms
    {2147483647, 5, 32, 129},   // tuned at n=1048576 k=1: 4.560 ms
    {2147483647, 52433, 64, 512},   // tuned at n=1048576 k=9: 6.151 ms
    {2147483647, 576716, 64, 33},   // tuned at n=1048576 k=104857: 99.909 ms
    {2147483647, 2147483647, 64, 33},   // tuned at n=1048576 k=1048576: 111.869 ms
};

// Ragged batches (one setting per batch), tuned on the 4096-instance ragged mix.
struct BatchTune {
    int B, sb_max;
};
static const BatchTune kBatchTune = {64, 0};   // 35.921 ms
