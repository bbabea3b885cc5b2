// Param.h -- scoring scheme and device counts of a search.
//
// The fields carry the reference's names and meaning (src/Param.h:59-117):
// gapOE = gapOpen + gapExtend, matrix[query code][db code], topNum hits.
// Configuration comes from the caller instead of argv.
#pragma once
#include <cstring>

class Param {
  public:
    Param(const int* plain32x32, int gapOpen, int gapExtend, int topNum, int hipNum,
          int cpuNum = 0)
        : cpuNum(cpuNum), hipNum(hipNum), gapOpen(gapOpen), gapExtend(gapExtend),
          gapOE(gapOpen + gapExtend), topNum(topNum) {
        for (int i = 0; i < 32; ++i)
            for (int j = 0; j < 32; ++j) {
                matrix[i][j] = plain32x32[i * 32 + j];
                plain_matrix[i * 32 + j] = plain32x32[i * 32 + j];
            }
    }
    int cpuNum;   // threads of the CPU worker (0 = none)
    int hipNum;   // HIPWorkers, devices 0..hipNum-1
    int gapOpen;
    int gapExtend;
    int gapOE;
    int topNum;
    int matrix[32][32];
    int plain_matrix[32 * 32];
};
