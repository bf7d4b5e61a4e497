// Text formats of the reference's command-line driver (textio.cpp).
#pragma once

#include <cstdint>
#include <string>

namespace fpp {

// Reads up to count floats, whitespace separated, in file order (fscanf "%f"
// semantics); out[count] is zeroed first, so values the file lacks, or values
// past the first token that is not a number, stay 0.  Returns the number read.
int64_t load_txt(const char* path, int64_t count, float* out);
// nq lines, each of the query's k ids followed by a space.
void write_topk_txt(const char* path, const int32_t* ids, int64_t nq, int k);
// Output file name of a verbose query: layout "1D" for query2Layers_1D
// (src/FalconnPP.cpp:694-700), "2D" for query2Layers (:1400-1406).
std::string topk_file_name(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                           int qProbes, const char* layout = "1D");

}  // namespace fpp
