// hss_host.h -- host-side helpers shared by the runtime's translation units
// (hss_runtime.hip, hss_compress.hip).  Not part of the public ABI.
#pragma once
#include <string>

#include "../../include/hss_hip.h"

namespace hss {

// sets the thread's last-error message and returns st
hss_status host_fail(hss_status st, const std::string& msg);
// the node table's tree structure (links, sizes, offsets; ranks and generator offsets
// are not looked at)
hss_status host_validate_tree(const hss_node_desc* nodes, int nn);
// a matrix over a generator blob already in device memory; on success the matrix owns
// *dblob and *dblob is set to null
hss_status host_mat_from_device(hss_ctx_t ctx, const hss_node_desc* nodes, int nn, double** dblob, int64_t blob_len,
                                hss_mat_t* out);
int host_ctx_device(hss_ctx_t ctx);
// a matrix's context, its node table (as created: ranks and blob offsets) and its
// generator blob in device memory
hss_ctx_t host_mat_ctx(hss_mat_t mat);
const hss_node_desc* host_mat_nodes(hss_mat_t mat, int* nn);
const double* host_mat_blob(hss_mat_t mat);
// rows of a factor's matrix
int64_t host_fact_rows(hss_fact_t fact);

}  // namespace hss
