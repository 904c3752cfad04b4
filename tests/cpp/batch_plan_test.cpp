// batch_plan_test.cpp -- host-only check of the batched launches' segment
// tables (rochpcg_amd/csrc/batch.hpp): the block assignment the planner
// uses, the blocks' segment search and the XCD block order, on segment lists
// read from stdin.  No HIP call is made; built with hipcc only because the
// header is HIP code.
//
// stdin: per launch a line "block vector_kind rows nseg" followed by nseg
// lines "off size".  For every launch it checks that
//   - the block map sends every logical block to the segment holding it;
//   - xcd_block is a permutation of the launch's blocks;
//   - every row (element) of every segment is visited by exactly one thread
//     (R consecutive rows per thread in row launches, chunks in vector ones);
//   - two rows per thread are used only where every offset and size is even.
// Prints "ok <launches> <blocks>" or the first failure.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../rochpcg_amd/csrc/batch.hpp"

using namespace hpcgx;

int main()
{
    long launches = 0, blocks_total = 0;
    int block, vector_kind, rows_req, nseg;
    while(scanf("%d %d %d %d", &block, &vector_kind, &rows_req, &nseg) == 4)
    {
        std::vector<BSeg> segs((size_t)nseg);
        for(BSeg& s : segs)
        {
            long long off, size;
            if(scanf("%lld %lld", &off, &size) != 2) return 2;
            s = BSeg{};
            s.off = off;
            s.size = size;
        }
        const int rows = vector_kind ? 1 : launch_rows(segs.data(), segs.size(), rows_req);
        bool even = true;
        for(const BSeg& s : segs) even = even && s.off % 2 == 0 && s.size % 2 == 0;
        if(rows == 2 && !even)
        {
            printf("FAIL launch %ld: two rows per thread on an odd segment\n", launches);
            return 1;
        }
        const int64_t nb = assign_blocks(segs.data(), segs.size(), block, vector_kind != 0, rows);
        // the block map against a linear scan; the XCD order is a permutation
        std::vector<char> table(table_bytes(segs.size(), nb));
        std::memcpy(table.data(), segs.data(), sizeof(BSeg) * segs.size());
        block_map(segs.data(), segs.size(), reinterpret_cast<uint16_t*>(table.data() + sizeof(BSeg) * segs.size()));
        const uint16_t* map = seg_map(reinterpret_cast<const BSeg*>(table.data()), nseg);
        std::vector<char> hit((size_t)nb, 0);
        for(int64_t bi = 0; bi < nb; ++bi)
        {
            const int64_t b = xcd_block(bi, nb);
            if(b < 0 || b >= nb || hit[(size_t)b])
            {
                printf("FAIL launch %ld: xcd_block is not a permutation at %lld\n", launches, (long long)bi);
                return 1;
            }
            hit[(size_t)b] = 1;
            int lin = -1;
            for(int j = 0; j < nseg; ++j)
                if(segs[(size_t)j].blk0 <= b && b < segs[(size_t)j].blk0 + segs[(size_t)j].nblk) lin = j;
            const int j = map[b];
            if(j != lin)
            {
                printf("FAIL launch %ld: block %lld in segment %d, the map gives %d\n", launches, (long long)b, lin, j);
                return 1;
            }
        }
        // coverage of every segment's rows, as the kernels index them
        for(const BSeg& s : segs)
        {
            std::vector<int> seen((size_t)s.size, 0);
            for(int64_t lb = 0; lb < s.nblk; ++lb)
            {
                if(vector_kind)
                {
                    const int64_t chunk = (s.size + s.nblk - 1) / s.nblk;
                    const int64_t e0 = lb * chunk, e1 = (e0 + chunk < s.size) ? e0 + chunk : s.size;
                    for(int64_t t = 0; t < block; ++t)
                        for(int64_t i = e0 + t; i < e1; i += block) seen[(size_t)i]++;
                }
                else
                {
                    for(int64_t t = 0; t < block; ++t)
                    {
                        const int64_t i = (lb * block + t) * rows;
                        if(i >= s.size) continue;
                        for(int k = 0; k < rows; ++k)
                        {
                            if(i + k >= s.size)
                            {
                                printf("FAIL launch %ld: row %lld past a segment of %lld\n", launches,
                                       (long long)(i + k), (long long)s.size);
                                return 1;
                            }
                            seen[(size_t)(i + k)]++;
                        }
                    }
                }
            }
            for(int64_t i = 0; i < s.size; ++i)
                if(seen[(size_t)i] != 1)
                {
                    printf("FAIL launch %ld: row %lld of a segment of %lld visited %d times\n", launches, (long long)i,
                           (long long)s.size, seen[(size_t)i]);
                    return 1;
                }
        }
        ++launches;
        blocks_total += nb;
    }
    printf("ok %ld %ld\n", launches, blocks_total);
    return 0;
}
