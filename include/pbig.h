/*
 * pbig.h -- lifecycle API of the reference (src/pbig/pbig.h:8-11), C++ linkage.
 *
 *   pbig_init(N, range, cs, deviceID, profSteps)
 *       N          maximum number of boxes per pbig_check call
 *       range      domain side: cells cover [0, range)^2 (coordinates outside
 *                  fold into the border cells)
 *       cs         cell side; 0 = the reference's automatic rule
 *                  (src/pbig/pbig.cu:1285-1357)
 *       deviceID   HIP device ordinal
 *       profSteps  accepted and ignored (the reference forces it to 0 in its
 *                  release build, src/pbig/pbig.cu:2147-2153)
 *   pbig_check(input, r)
 *       input      N RectInfo records with rid == index
 *       r          result sink; r(rid1, rid2) is called once per pair
 *   pbig_finalize()
 *
 * Errors print "ERROR:..." and exit(-1), as the reference's CUDA_ERR does
 * (src/pbig/log.h:4-13).  The C-ABI in pbig_hip.h returns status codes
 * instead and is the interface to bind from other languages.
 */
#ifndef __PBIG_H_
#define __PBIG_H_

#include <set>
#include <vector>

struct report;
struct RectInfo;
void pbig_init(unsigned int, unsigned int, unsigned int, unsigned int, unsigned int);
void pbig_check(struct RectInfo *, struct report *);
void pbig_finalize();

/* Extensions (no counterpart in the reference's pbig.h): the pairs of a
 * check as the reference's compressed record stream (64-bit words in the
 * layout of src/pbig/pbig.cu:849-888) with the cell arrays a reader needs,
 * and a decoder that calls the report once per pair (decode.cc:26-138, with
 * the partner offsets read where the encoder writes them). */
void pbig_check_stream(struct RectInfo *input, std::vector<unsigned long long> *words,
                       std::vector<unsigned int> *cell_start, std::vector<unsigned int> *cell_recs);
void pbig_decode_stream(const std::vector<unsigned long long> &words, const std::vector<unsigned int> &cell_start,
                        const std::vector<unsigned int> &cell_recs, struct report *r);

#endif
