/**
 * @file cmpconst.h
 * @brief Tolerance of compareBrick (include/cmpconst.h:9 of bricklib).
 */
#ifndef BK_CMPCONST_H
#define BK_CMPCONST_H

#ifndef BRICK_TOLERANCE
#define BRICK_TOLERANCE 1e-6
#endif

#endif  // BK_CMPCONST_H
