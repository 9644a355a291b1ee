"""Work and roofline constants of the BSSN RHS + RK4 path on MI355X.

FLOPS_PER_PT_RHS is the operation count of the reference algorithm as
written (bssn_rhs.f90 with its stencil routines, GAUGE 0) for one RHS
evaluation at an interior point: add/sub/mul/div one each, counted by
running the C restatement (bit-identical to the reference) with a counting
scalar type (oracle/flopcount.cpp): 4394 add/sub + 2826 mul + 376 div.
The predictor's co == 0 call adds the constraint residuals (8492 total).

BYTES_PER_PT_STAGE are the compulsory HBM bytes of the fused RHS + RK4
kernel per interior point and RK4 sub-step: stage 0 reads the state and
writes the accumulator and the stage output (72 FP64 fields); stages 1-2
read the stage input, base state and accumulator and write accumulator and
output (120); stage 3 reads input, base and accumulator and writes the new
state (96).  Halo re-reads are not compulsory and not counted.  A
predictor that forms the constraint residuals (the reference's co = 0 call)
also writes the 7 residual fields: BYTES_PER_PT_STAGE_RES.

BYTES_PER_PT_GHOST_FILL are the compulsory bytes of the frozen-ghost fill
per ghost point (read the 24 base-state fields, write the 24 outputs).  On
a level patch with 3 ghost planes and plain launches the fill runs in the
same launch as sub-steps 0 and 1 (sub-step 2 writes the buffer sub-step 0
filled; sub-step 3's fill is a launch of its own), over the
(n + 6)^3 - n^3 ghost points: GHOST_FILL_STAGES.
"""
FLOPS_PER_PT_RHS = 7596
FLOPS_PER_PT_RHS_CO0 = 8492
BYTES_PER_PT_STAGE = (72 * 8, 120 * 8, 120 * 8, 96 * 8)
BYTES_PER_PT_STAGE_RES = (79 * 8, 120 * 8, 120 * 8, 96 * 8)
FLOPS_PER_PT_STAGE_RES = (FLOPS_PER_PT_RHS_CO0, FLOPS_PER_PT_RHS, FLOPS_PER_PT_RHS, FLOPS_PER_PT_RHS)
BYTES_PER_PT_GHOST_FILL = 48 * 8
GHOST_FILL_STAGES = (1, 1, 0, 0)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0        # HBM3E, GB/s
PEAK_FP64_TFLOPS = 78.6      # FP64 vector (= FP64 matrix on CDNA4), TFLOP/s
