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
    {2560, 5, 32, 512},   // tuned at n=1024 k=1: 0.168 ms
    {2560, 55, 32, 512},   // tuned at n=1024 k=9: 0.180 ms
    {2560, 563, 64, 257},   // tuned at n=1024 k=102: 0.219 ms
    {2560, 2147483647, 32, 512},   // tuned at n=1024 k=1024: 0.264 ms
    {10240, 5, 32, 65},   // tuned at n=4096 k=1: 0.178 ms
    {10240, 209, 64, 65},   // tuned at n=4096 k=9: 0.208 ms
    {10240, 2252, 64, 65},   // tuned at n=4096 k=409: 0.470 ms
    {10240, 2147483647, 64, 65},   // tuned at n=4096 k=4096: 0.497 ms
    {40960, 5, 64, 512},   // tuned at n=16384 k=1: 0.212 ms
    {40960, 823, 64, 512},   // tuned at n=16384 k=9: 0.260 ms
    {40960, 9011, 64, 65},   // tuned at n=16384 k=1638: 0.986 ms
    {40960, 2147483647, 64, 65},   // tuned at n=16384 k=16384: 1.368 ms
    {163840, 5, 32, 129},   // tuned at n=65536 k=1: 0.382 ms
    {163840, 3281, 64, 257},   // tuned at n=65536 k=9: 0.500 ms
    {163840, 36044, 64, 129},   // tuned at n=65536 k=6553: 5.867 ms
    {163840, 2147483647, 64, 129},   // tuned at n=65536 k=65536: 6.959 ms
    {655360, 5, 32, 257},   // tuned at n=262144 k=1: 1.347 ms
    {655360, 13111, 64, 257},   // tuned at n=262144 k=9: 1.765 ms
    {655360, 144179, 64, 129},   // tuned at n=262144 k=26214: 30.141 ms
    {655360, 2147483647, 64, 129},   // tuned at n=262144 k=262144: 34.491 ms
    {2147483647, 5, 32, 512},   // tuned at n=1048576 k=1: 5.230 ms
    {2147483647, 52433, 64, 129},   // tuned at n=1048576 k=9: 6.806 ms
    {2147483647, 576716, 64, 129},   // tuned at n=1048576 k=104857: 150.577 ms
    {2147483647, 2147483647, 64, 129},   // tuned at n=1048576 k=1048576: 168.935 ms
};
