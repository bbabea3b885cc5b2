// dbmaker -- FASTA protein database -> <out_base>.seq/.map/.title, the files
// the reference's DBmaker writes (src/DBmaker.cpp:9-28), for swhybrid -db.
#include <cstdio>

#include "swhybrid_host.h"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: dbmaker <database.fasta> <out_base>\n");
        return 2;
    }
    std::printf("Processing database %s\nTarget file basename %s\n", argv[1], argv[2]);
    uint64_t records = 0, slots = 0;
    if (swhy_dbfile_make(argv[1], argv[2], 0, &records, &slots) != 0) {
        std::fprintf(stderr, "dbmaker: %s\n", swhy_last_error());
        return 1;
    }
    std::printf("total sequence num is %llu (%llu with the final batch's fillers)\n",
                (unsigned long long)records, (unsigned long long)slots);
    return 0;
}
