/*
 * sdpa_io.c -- reader of SDPA sparse-format files into the solver's input layout.
 *
 * Restates HReadSDPA (interface/hdsdp_file_io.c:34-381): the problem
 *     min <F0, X>  s.t.  <F_i, X> = c_i  (SDPA: max c^T y, F0 - sum y_i F_i >= 0 dual)
 * is handed to the solver as C = -F0 (the reference negates the F0 entries,
 * interface/hdsdp_file_io.c:234-236, 248-250), A_i = F_i and b = c.  Each SDP block
 * becomes one CSC matrix of m + 1 columns over the packed-lower row index
 * (2 n - j - 1) j / 2 + i of entry (i, j), i >= j (column 0 is C); a negative last block
 * size is a diagonal (LP) block, returned as one CSC of m + 1 columns over its rows.
 * Entries of magnitude below 1e-12 are dropped, as in the reference; entries are kept in
 * file order within a column (duplicates are not summed), which is what the reference's
 * triplet compression produces.
 *
 * Parsing: comment lines start with '*' or '"'; the block-size and right-hand-side lines
 * may use the separators { } ( ) , and spaces; entry lines are
 * "constraint block row column value" with 1-based block, row and column (constraint 0
 * is F0), read until the end of the file or a "BEGIN.COMMENT" line.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsdp_host.h"
#include "hdsdp_file_io.h"

#define SDPA_LINE 4096

typedef struct {
    int col;
    int row;
    double val;
} sdpa_entry;

typedef struct {
    sdpa_entry *e;
    long n, cap;
} sdpa_list;

static int list_push( sdpa_list *l, int row, int col, double val ) {
    if ( l->n == l->cap ) {
        long cap = l->cap ? 2 * l->cap : 1024;
        sdpa_entry *p = (sdpa_entry *) realloc(l->e, sizeof(sdpa_entry) * (size_t) cap);
        if ( !p ) return 0;
        l->e = p;
        l->cap = cap;
    }
    l->e[l->n].row = row;
    l->e[l->n].col = col;
    l->e[l->n].val = val;
    l->n += 1;
    return 1;
}

/* CSC with nCol columns from the entry list, entries of a column in list order. */
static hdsdp_retcode list_to_csc( const sdpa_list *l, int nCol, int **pBeg, int **pIdx, double **pElem ) {
    int *beg = (int *) calloc((size_t) nCol + 1, sizeof(int));
    int *idx = (int *) malloc(sizeof(int) * (size_t) (l->n ? l->n : 1));
    double *elem = (double *) malloc(sizeof(double) * (size_t) (l->n ? l->n : 1));
    int *next = (int *) malloc(sizeof(int) * (size_t) nCol + sizeof(int));
    if ( !beg || !idx || !elem || !next ) {
        free(beg);
        free(idx);
        free(elem);
        free(next);
        return HDSDP_RETCODE_MEMORY;
    }
    for ( long k = 0; k < l->n; ++k ) beg[l->e[k].col + 1] += 1;
    for ( int j = 0; j < nCol; ++j ) beg[j + 1] += beg[j];
    memcpy(next, beg, sizeof(int) * (size_t) nCol);
    for ( long k = 0; k < l->n; ++k ) {
        const int q = next[l->e[k].col]++;
        idx[q] = l->e[k].row;
        elem[q] = l->e[k].val;
    }
    free(next);
    *pBeg = beg;
    *pIdx = idx;
    *pElem = elem;
    return HDSDP_RETCODE_OK;
}

/* Next line that is not a comment; 0 at end of file. */
static int next_data_line( FILE *f, char *line ) {
    while ( fgets(line, SDPA_LINE, f) ) {
        if ( line[0] != '*' && line[0] != '"' ) return 1;
    }
    return 0;
}

/* Replace the SDPA separators by blanks. */
static void blank_separators( char *s ) {
    for ( ; *s; ++s )
        if ( *s == '{' || *s == '}' || *s == '(' || *s == ')' || *s == ',' || *s == '\'' ) *s = ' ';
}

/* Read `count` numbers from the rest of the file's lines (separators allowed). */
static int read_numbers( FILE *f, char *line, int count, double *out ) {
    int got = 0;
    while ( got < count ) {
        if ( !fgets(line, SDPA_LINE, f) ) return got;
        blank_separators(line);
        char *p = line, *end = NULL;
        while ( got < count ) {
            while ( *p && isspace((unsigned char) *p) ) ++p;
            if ( !*p ) break;
            double v = strtod(p, &end);
            if ( end == p ) return -1;
            out[got++] = v;
            p = end;
        }
    }
    return got;
}

hdsdp_retcode HReadSDPA( char *fname, int *pnConstrs, int *pnBlks, int **pblkDims, double **prowRHS,
                         int ***pconeMatBeg, int ***pconeMatIdx, double ***pconeMatElem, int *pnCols,
                         int *pnLPCols, int **pLpMatBeg, int **pLpMatIdx, double **pLpMatElem, int *pnElems ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    char *line = (char *) malloc(SDPA_LINE);
    FILE *f = NULL;
    int nConstrs = 0, nBlocks = 0, nSdp = 0, nLP = 0, nCols = 0, nElems = 0, warned = 0;
    int *dims = NULL, **cBeg = NULL, **cIdx = NULL, *lBeg = NULL, *lIdx = NULL;
    double *rhs = NULL, **cElem = NULL, *lElem = NULL, *rawDims = NULL;
    sdpa_list *lists = NULL, lpList = {NULL, 0, 0};

    if ( !line ) return HDSDP_RETCODE_MEMORY;
    f = fopen(fname, "r");
    if ( !f ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    /* number of constraints, number of blocks */
    if ( !next_data_line(f, line) || sscanf(line, "%d", &nConstrs) != 1 || nConstrs <= 0 ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    if ( !next_data_line(f, line) || sscanf(line, "%d", &nBlocks) != 1 || nBlocks <= 0 ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    /* block sizes: positive SDP blocks, the last one may be a negative (diagonal) block */
    rawDims = (double *) malloc(sizeof(double) * (size_t) nBlocks);
    dims = (int *) calloc((size_t) nBlocks, sizeof(int));
    rhs = (double *) malloc(sizeof(double) * (size_t) nConstrs);
    if ( !rawDims || !dims || !rhs ) {
        retcode = HDSDP_RETCODE_MEMORY;
        goto exit_cleanup;
    }
    if ( read_numbers(f, line, nBlocks, rawDims) != nBlocks ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    for ( int k = 0; k < nBlocks; ++k ) {
        const int d = (int) rawDims[k];
        if ( d > 0 ) {
            dims[nSdp++] = d;
            nCols += d * d;
        } else if ( d < 0 && k == nBlocks - 1 ) {
            nLP = -d;
        } else {
            /* only the last block may be diagonal (interface/hdsdp_file_io.c:107-113) */
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
    }
    if ( read_numbers(f, line, nConstrs, rhs) != nConstrs ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    lists = (sdpa_list *) calloc((size_t) (nSdp ? nSdp : 1), sizeof(sdpa_list));
    if ( !lists ) {
        retcode = HDSDP_RETCODE_MEMORY;
        goto exit_cleanup;
    }
    /* entries */
    while ( fgets(line, SDPA_LINE, f) ) {
        int iCon, iBlk, iRow, iCol;
        double v;
        if ( line[0] == '*' || line[0] == '"' ) continue;
        if ( sscanf(line, "%d %d %d %d %lg", &iCon, &iBlk, &iRow, &iCol, &v) != 5 ) {
            if ( strncmp(line, "BEGIN.COMMENT", 13) == 0 ) break;
            char *p = line;
            while ( *p && isspace((unsigned char) *p) ) ++p;
            if ( !*p ) continue; /* blank line */
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
        iBlk -= 1;
        iRow -= 1;
        iCol -= 1;
        if ( iCon < 0 || iCon > nConstrs || iBlk < 0 || iBlk >= nSdp + (nLP > 0) || iRow < 0 || iCol < 0 ) {
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
        if ( fabs(v) < 1e-12 ) {
            if ( !warned ) {
                printf("[Warning] Entry smaller than 1e-12 is ignored. \n");
                warned = 1;
            }
            continue;
        }
        if ( iCon == 0 ) v = -v; /* C = -F0 */
        int ok;
        if ( iBlk == nSdp ) {
            if ( iRow >= nLP ) {
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
            ok = list_push(&lpList, iRow, iCon, v);
        } else {
            const int n = dims[iBlk];
            if ( iRow >= n || iCol >= n ) {
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
            if ( iRow > iCol ) {
                const int t = iRow;
                iRow = iCol;
                iCol = t;
            }
            /* lower-triangle entry (iCol, iRow), packed by columns */
            const long pk = (long) (2 * n - iRow - 1) * iRow / 2 + iCol;
            ok = list_push(&lists[iBlk], (int) pk, iCon, v);
        }
        if ( !ok ) {
            retcode = HDSDP_RETCODE_MEMORY;
            goto exit_cleanup;
        }
        nElems += 1;
    }
    cBeg = (int **) calloc((size_t) (nSdp ? nSdp : 1), sizeof(int *));
    cIdx = (int **) calloc((size_t) (nSdp ? nSdp : 1), sizeof(int *));
    cElem = (double **) calloc((size_t) (nSdp ? nSdp : 1), sizeof(double *));
    if ( !cBeg || !cIdx || !cElem ) {
        retcode = HDSDP_RETCODE_MEMORY;
        goto exit_cleanup;
    }
    for ( int k = 0; k < nSdp; ++k )
        if ( (retcode = list_to_csc(&lists[k], nConstrs + 1, &cBeg[k], &cIdx[k], &cElem[k])) != HDSDP_RETCODE_OK )
            goto exit_cleanup;
    if ( nLP > 0 && (retcode = list_to_csc(&lpList, nConstrs + 1, &lBeg, &lIdx, &lElem)) != HDSDP_RETCODE_OK )
        goto exit_cleanup;

    *pnConstrs = nConstrs;
    *pnBlks = nSdp;
    *pblkDims = dims;
    *prowRHS = rhs;
    *pconeMatBeg = cBeg;
    *pconeMatIdx = cIdx;
    *pconeMatElem = cElem;
    *pnCols = nCols;
    *pnLPCols = nLP;
    *pLpMatBeg = lBeg;
    *pLpMatIdx = lIdx;
    *pLpMatElem = lElem;
    *pnElems = nElems;

exit_cleanup:
    if ( f ) fclose(f);
    free(line);
    free(rawDims);
    if ( lists ) {
        for ( int k = 0; k < nSdp; ++k ) free(lists[k].e);
        free(lists);
    }
    free(lpList.e);
    if ( retcode != HDSDP_RETCODE_OK ) {
        free(dims);
        free(rhs);
        free(lBeg);
        free(lIdx);
        free(lElem);
        for ( int k = 0; k < nSdp; ++k ) {
            if ( cBeg ) free(cBeg[k]);
            if ( cIdx ) free(cIdx[k]);
            if ( cElem ) free(cElem[k]);
        }
        free(cBeg);
        free(cIdx);
        free(cElem);
    }
    return retcode;
}

void HFreeSDPA( int nBlks, int *blkDims, double *rowRHS, int **coneMatBeg, int **coneMatIdx, double **coneMatElem,
                int *LpMatBeg, int *LpMatIdx, double *LpMatElem ) {
    for ( int k = 0; k < nBlks; ++k ) {
        if ( coneMatBeg ) free(coneMatBeg[k]);
        if ( coneMatIdx ) free(coneMatIdx[k]);
        if ( coneMatElem ) free(coneMatElem[k]);
    }
    free(coneMatBeg);
    free(coneMatIdx);
    free(coneMatElem);
    free(blkDims);
    free(rowRHS);
    free(LpMatBeg);
    free(LpMatIdx);
    free(LpMatElem);
}
