/// C entry points for reference-data ("PDMC") files, for callers without the C++ header
/// (Python ctypes, Julia).  They wrap pdmk/campaign_ref_io.hpp, which keeps the layout of
/// the reference's cpp/include/pdmk/campaign_ref_io.hpp:12-186 (header-only C++ there, so the
/// reference has no C symbols for it).  Host only; return 0 on success, -1 on error with a
/// message on stderr.
#ifndef PDMK_CAMPAIGN_H
#define PDMK_CAMPAIGN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/// SHA-256 of data[0 .. len) into out32[32] (pdmk::campaign::sha256; reference
/// campaign_ref_io.hpp:58-111).
int pdmk_campaign_sha256(const unsigned char* data, size_t len, unsigned char* out32);

/// Writes a PDMC file (pdmk::campaign::write_ref; reference campaign_ref_io.hpp:131-150).
/// cell9 column-major, pos_3xN column-major, ref_idx / ref_u of length M, chg of length N.
int pdmk_campaign_write(const char* path, unsigned config_id, unsigned N, int digits, int prec,
                        unsigned long long seed, const double* cell9, unsigned M, const unsigned* ref_idx,
                        const double* ref_u, const double* pos_3xN, const double* chg);

/// Reads and verifies a PDMC file's header fields (any pointer may be NULL).  The body is
/// verified too: a file that fails read_ref fails here.
int pdmk_campaign_read_header(const char* path, unsigned* config_id, unsigned* N, int* digits, int* prec,
                              unsigned long long* seed, unsigned* M);

/// Reads and verifies a PDMC file (pdmk::campaign::read_ref; reference
/// campaign_ref_io.hpp:152-184) into caller arrays sized from pdmk_campaign_read_header.
int pdmk_campaign_read(const char* path, double* cell9, unsigned* ref_idx, double* ref_u, double* pos_3xN,
                       double* chg);

#ifdef __cplusplus
}
#endif

#endif  // PDMK_CAMPAIGN_H
