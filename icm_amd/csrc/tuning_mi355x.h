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
    {2560, 5, 32, 512},   // tuned at n=1024 k=1: 0.163 ms
    {2560, 55, 32, 257},   // tuned at n=1024 k=9: 0.175 ms
    {2560, 563, 64, 129},   // tuned at n=1024 k=102: 0.213 ms
    {2560, 2147483647, 32, 512},   // tuned at n=1024 k=1024: 0.258 ms
    {10240, 5, 32, 129},   // tuned at n=4096 k=1: 0.171 ms
    {10240, 209, 64, 512},   // tuned at n=4096 k=9: 0.202 ms
    {10240, 2252, 64, 65},   // tuned at n=4096 k=409: 0.468 ms
    {10240, 2147483647, 64, 65},   // tuned at n=4096 k=4096: 0.492 ms
    {40960, 5, 64, 257},   // tuned at n=16384 k=1: 0.201 ms
    {40960, 823, 64, 512},   // tuned at n=16384 k=9: 0.246 ms
    {40960, 9011, 64, 33},   // tuned at n=16384 k=1638: 0.775 ms
    {40960, 2147483647, 64, 33},   // tuned at n=16384 k=16384: 1.006 ms
    {163840, 5, 32, 512},   // tuned at n=65536 k=1: 0.351 ms
    {163840, 3281, 64, 512},   // tuned at n=65536 k=9: 0.483 ms
    {163840, 36044, 64, 33},   // tuned at n=65536 k=6553: 3.995 ms
    {163840, 2147483647, 64, 33},   // tuned at n=65536 k=65536: 4.775 ms
    {655360, 5, 32, 129},   // tuned at n=262144 k=1: 1.184 ms
    {655360, 13111, 64, 65},   // tuned at n=262144 k=9: 1.595 ms
    {655360, 144179, 64, 33},   // tuned at n=262144 k=26214: 19.502 ms
    {655360, 2147483647, 64, 33},   // tuned at n=262144 k=262144: 22.393 ms
    {2147483647, 5, 32, 129},   // tuned at n=1048576 k=1: 4.482 ms
    {2147483647, 52433, 64, 512},   // tuned at n=1048576 k=9: 6.104 ms
    {2147483647, 576716, 64, 0},   // tuned at n=1048576 k=104857: 93.617 ms
    {2147483647, 2147483647, 64, 17},   // tuned at n=1048576 k=1048576: 105.140 ms
};

// Ragged batches (one setting per batch), tuned on the 4096-instance ragged mix.
struct BatchTune {
    int B, sb_max;
};
static const BatchTune kBatchTune = {64, 0};   // 36.069 ms
