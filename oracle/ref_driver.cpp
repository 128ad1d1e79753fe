// ORACLE (reference side) -- test infrastructure, not product code.
//
// Compiles the reference's own scalar C++ LOMS port, bench/cpp/loms_decode.cpp,
// unchanged and from where it lies (the include path points into the
// reference tree; nothing is copied), and exposes its fixed-configuration
// decode (BG1, Z=384, beta=0.25, exactly 10 iterations, no early exit,
// loms_decode.cpp:28-36, 285-321) through a C entry point, so the oracle's
// f32 restatement can be compared with it value for value.
//
// Built by oracle/Makefile into oracle/_ref/ only where the reference tree
// is mounted; tests skip the comparison when it is absent.
#define main loms_reference_program_main
#include "bench/cpp/loms_decode.cpp"
#undef main

extern "C" int ref_loms_bg1_z384_decode10(const char *json_text, float *llr, float *edge_r)
{
    std::vector<BgEntry> entries = parse_bg1_entries(json_text);
    if ((int)entries.size() != TOTAL_EDGES) return -1;
    DecoderLayout L = build_layout(entries);
    std::vector<float> q((size_t)L.max_layer_degree * Z);
    decode_loms(llr, edge_r, q.data(), L);
    return 0;
}

extern "C" uint64_t ref_loms_checksum(const float *llr, int n) { return hard_decision_checksum(llr, n); }
