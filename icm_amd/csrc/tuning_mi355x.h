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
    {2048, 3, 32, 512},   // tuned at n=1024 k=1: 0.167 ms
    {2048, 28, 32, 512},   // tuned at n=1024 k=9: 0.177 ms
    {2048, 95, 32, 257},   // tuned at n=1024 k=90: 0.211 ms
    {2048, 302, 64, 257},   // tuned at n=1024 k=102: 0.213 ms
    {2048, 960, 64, 257},   // tuned at n=1024 k=900: 0.256 ms
    {2048, 2147483647, 64, 257},   // tuned at n=1024 k=1024: 0.254 ms
    {8192, 3, 32, 65},   // tuned at n=4096 k=1: 0.169 ms
    {8192, 28, 64, 257},   // tuned at n=4096 k=9: 0.209 ms
    {8192, 191, 64, 129},   // tuned at n=4096 k=90: 0.340 ms
    {8192, 606, 64, 33},   // tuned at n=4096 k=409: 0.455 ms
    {8192, 1920, 64, 65},   // tuned at n=4096 k=900: 0.463 ms
    {8192, 2147483647, 64, 17},   // tuned at n=4096 k=4096: 0.473 ms
    {32768, 3, 32, 512},   // tuned at n=16384 k=1: 0.190 ms
    {32768, 28, 64, 257},   // tuned at n=16384 k=9: 0.241 ms
    {32768, 284, 64, 512},   // tuned at n=16384 k=90: 0.502 ms
    {32768, 1214, 64, 33},   // tuned at n=16384 k=900: 0.687 ms
    {32768, 5180, 64, 33},   // tuned at n=16384 k=1638: 0.737 ms
    {32768, 2147483647, 64, 33},   // tuned at n=16384 k=16384: 0.908 ms
    {131072, 3, 32, 512},   // tuned at n=65536 k=1: 0.345 ms
    {131072, 28, 64, 129},   // tuned at n=65536 k=9: 0.469 ms
    {131072, 284, 64, 33},   // tuned at n=65536 k=90: 1.313 ms
    {131072, 2428, 64, 17},   // tuned at n=65536 k=900: 2.229 ms
    {131072, 20723, 64, 17},   // tuned at n=65536 k=6553: 3.713 ms
    {131072, 2147483647, 64, 17},   // tuned at n=65536 k=65536: 4.433 ms
    {524288, 3, 32, 257},   // tuned at n=262144 k=1: 1.175 ms
    {524288, 28, 64, 257},   // tuned at n=262144 k=9: 1.599 ms
    {524288, 284, 64, 33},   // tuned at n=262144 k=90: 4.511 ms
    {524288, 4857, 64, 33},   // tuned at n=262144 k=900: 8.342 ms
    {524288, 82896, 64, 33},   // tuned at n=262144 k=26214: 18.199 ms
    {524288, 2147483647, 64, 0},   // tuned at n=262144 k=262144: 20.805 ms
    {2147483647, 3, 32, 257},   // tuned at n=1048576 k=1: 4.468 ms
    {2147483647, 28, 64, 65},   // tuned at n=1048576 k=9: 6.043 ms
    {2147483647, 284, 64, 0},   // tuned at n=1048576 k=90: 17.094 ms
    {2147483647, 9714, 64, 33},   // tuned at n=1048576 k=900: 32.447 ms
    {2147483647, 331587, 64, 0},   // tuned at n=1048576 k=104857: 87.216 ms
    {2147483647, 2147483647, 64, 33},   // tuned at n=1048576 k=1048576: 98.883 ms
};

// Ragged batches (one setting per batch), tuned on the 4096-instance ragged mix.
struct BatchTune {
    int B, sb_max;
};
static const BatchTune kBatchTune = {64, 0};   // 35.522 ms

// Small instances in one launch (k_small): used when the tree fits one
// workgroup's LDS and its schoolbook multiply-adds per quadrature node are at
// most this many (crossover against the multi-launch path).
static const double kSmallMaxWork = 400000.0;   // n=1024: k=100 (464k) ties, k=30 (158k) 1.6x faster
