/*
 * hdsdp_file_io.h -- SDPA sparse-format input.
 *
 * HReadSDPA has the reference's signature and output layout
 * (interface/hdsdp_file_io.h, interface/hdsdp_file_io.c:34-381): per SDP block a CSC
 * matrix of nConstrs + 1 columns over packed-lower row indices, column 0 holding
 * C = -F0 and column i holding A_i = F_i, ready for HUserDataSetConeData; the right-hand
 * side b = c; a negative last block size is returned as the LP block.  All outputs are
 * allocated with malloc.
 */
#ifndef HDSDP_AMD_HDSDP_FILE_IO_H
#define HDSDP_AMD_HDSDP_FILE_IO_H

#include "hdsdp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* interface/hdsdp_file_io.h */
hdsdp_retcode HReadSDPA( char *fname, int *pnConstrs, int *pnBlks, int **pblkDims, double **prowRHS,
                         int ***pconeMatBeg, int ***pconeMatIdx, double ***pconeMatElem, int *pnCols,
                         int *pnLPCols, int **pLpMatBeg, int **pLpMatIdx, double **pLpMatElem, int *pnElems );

/* Extension (not in the reference): frees what HReadSDPA returned. */
void HFreeSDPA( int nBlks, int *blkDims, double *rowRHS, int **coneMatBeg, int **coneMatIdx, double **coneMatElem,
                int *LpMatBeg, int *LpMatIdx, double *LpMatElem );

#ifdef __cplusplus
}
#endif

#endif /* HDSDP_AMD_HDSDP_FILE_IO_H */
