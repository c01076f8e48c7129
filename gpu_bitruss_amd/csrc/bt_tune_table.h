// Counting launch settings per graph-size class, measured on MI355X by
// gpu_bitruss_amd/tune.py (which rewrites this file). Class c holds graphs
// of 4^c to 4^(c+1) - 1 edges (batch unions: edges per graph). Fields:
// wave_wedges, block_wedges, block_threads, hub_threads.
#pragma once
static const bt_count_config kTuneTable[BT_TUNE_CLASSES] = {
    {256, 4096, 512, 512},  // class 0: as class 3
    {256, 4096, 512, 512},  // class 1: as class 3
    {256, 4096, 512, 512},  // class 2: as class 3
    {256, 4096, 512, 512},  // class 3: mix_graph 84, m 121: 0.034 ms (previous 0.035 ms)
    {256, 16384, 512, 256},  // class 4: mix_graph 35, m 289: 0.034 ms (previous 0.036 ms)
    {256, 16384, 512, 256},  // class 5: mix_graph 0, m 2708: 0.036 ms (previous 0.037 ms)
    {256, 8192, 512, 512},  // class 6: mix_graph 28, m 8852: 0.044 ms (previous 0.045 ms)
    {512, 4096, 512, 512},  // class 7: mix_graph 42, m 42164: 0.100 ms (previous 0.101 ms)
    {256, 8192, 512, 256},  // class 8: mix_graph 196, m 204551: 0.442 ms (previous 0.478 ms)
    {256, 16384, 512, 256},  // class 9: mix_graph 14, m 290878: 0.651 ms (previous 0.940 ms)
    {1024, 16384, 512, 512},  // class 10: powerlaw_2m, m 1981225: 8.232 ms (previous 11.066 ms)
    {1024, 16384, 512, 512},  // class 11: as class 10
    {1024, 16384, 512, 512},  // class 12: skewed_60m, m 58883633: 1079.367 ms (previous 1220.452 ms)
    {1024, 16384, 512, 512},  // class 13: as class 12
    {1024, 16384, 512, 512},  // class 14: as class 12
    {1024, 16384, 512, 512},  // class 15: as class 12
};
