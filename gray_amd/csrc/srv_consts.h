/*
 * srv_consts.h -- sizes shared by the serving pool's device kernels
 * (serve.h) and its host scheduler (srv_sched.h, which includes no HIP).
 */
#pragma once

namespace gray {

enum { SRV_KERR = 0, SRV_GRID = 1 };
enum { SRV_NF = 4 };                  /* channels a request may carry           */
enum { SRV_MAX_CLASSES = 8 };         /* Kerr classes served at once            */
enum { SRV_BLOCK = 256 };             /* threads per Kerr / FP64 block           */
enum { SRV_GWAVES = 4 };              /* waves per grid block                    */

} // namespace gray
