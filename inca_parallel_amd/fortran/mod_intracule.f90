!! calculate_intracule: the drop-in for the reference's hot path
!! (src/mod_intracule.f90:18-572).  Same name, argument and globals; the
!! quartet loop, the grids and the grid reductions run on the GPU through
!! libinca_hip.so, and this module keeps the reference's output files:
!! the .out report (:410-505), the .rad radial scan (:512-533), the .cube
!! vector intracule (:552-568) and <stem>_smearing.dat (:535-550, 576-649).
module intracule
   use iso_c_binding
   use geninfo
   use numbers
   use intrainfo
   use quadratures
   use inca_hip
   use autogrid, only: centercalc
   implicit none

   integer(c_int64_t), private :: counts(4) = 0
   double precision, private :: kern_ms(3) = 0.d0
   integer(c_int64_t), private :: evals_ref = 0, evals_exec = 0
   integer(c_int64_t), private :: ngroups = 0, ngroups64 = 0, ncoef = 0
   character(len=200), private :: grid_note = ''

contains

   subroutine calculate_intracule(normalize_dm2p)
      logical, intent(in) :: normalize_dm2p
      integer, parameter :: iout = 15
      integer(c_int64_t) :: nrec, nr, ng
      integer(c_int32_t) :: fmt
      integer(c_int) :: rc
      integer(c_int32_t), allocatable, target :: tmn32(:)
      double precision, allocatable, target :: rintra(:), ivec(:), ipc(:), vpc(:)
      double precision, allocatable, target :: pts(:,:), su(:), sp(:), sia(:), radl(:)
      integer, parameter :: n_bins = 1000
      double precision, target :: rintegral, vee
      double precision :: lim, t1, t2, t3, t4, gridt, u_max, ivz(1), iv2(2)
      integer :: i, npairs
      integer(c_int64_t) :: refval, np2

      open(unit=iout, file=outname)
      t1 = wall()
      call inca_open()
      allocate(tmn32(3*nprim))
      tmn32 = reshape(int(TMN(1:nprim,1:3), c_int32_t), (/3*nprim/))
      if (.not. allocated(Xn)) then
         ! no .bas file: primitive centres from the atom they sit on
         allocate(Xn(nprim), Yn(nprim), Zn(nprim))
         do i = 1, nprim
            Xn(i) = cartes(Ra(i),1); Yn(i) = cartes(Ra(i),2); Zn(i) = cartes(Ra(i),3)
         end do
      end if
      call inca_check(inca_set_primitives(inca_ctx, int(nprim, c_int32_t), Alpha, tmn32, Xn, Yn, &
                      Zn, merge(1_c_int32_t, 0_c_int32_t, normalize_dm2p)), "inca_set_primitives")
      maxgrid = 0; rgrid = 0; rrgrid = 0

      ! grid of the main job (mod_intracule.f90:91-112)
      if (radial_integral) then
         call multicenter_grid()
         write(*,*) "Number of grid points to compute intracule:", rrgrid
      else if (radial_plot) then
         call inca_check(inca_grid_radial_plot(inca_ctx, int(nblock, c_int32_t), tart, stp, &
              int(n_an_per_part, c_int32_t), merge(1_c_int32_t, 0_c_int32_t, nosym), nr, ng), &
              "inca_grid_radial_plot")
         nradi = int(nr); rgrid = int(ng)
         allocate(radl(max(nradi,1)))
         call inca_check(inca_get_radial_layout(inca_ctx, c_loc(radl), c_null_ptr), &
              "inca_get_radial_layout")
         if (allocated(radi)) deallocate(radi)
         allocate(radi(nradi))
         radi = radl(1:nradi)
         write(*,*) "Number of grid points to compute intracule:", rgrid
      else if (cubeintra) then
         call inca_check(inca_grid_cube(inca_ctx, center_i, step_i, int(np_i, c_int32_t), ng), &
              "inca_grid_cube")
         rgrid = int(ng)
         write(*,*) "Number of grid points to compute intracule:", rgrid
      end if
      t2 = wall()
      gridt = t2 - t1

      ! DM2 stream (mod_intracule.f90:164-195)
      rc = inca_load_dm2p(inca_ctx, c_string(dm2name), nrec, fmt)
      if (rc == -4) then
         write(*,*) "Error: Unsupported file format with markers. Please convert to stream format."
         stop
      end if
      call inca_check(rc, "inca_load_dm2p")
      if (fmt == 1) then
         write(*,*) "INFO: File Format is Standard Fortran (24-byte markers)."
      else
         write(*,*) "INFO: File Format appears to be Flat Stream (No markers)."
      end if
      if (mixed_precision) call inca_check(inca_set_mixed_tail(inca_ctx, mixed_tail), &
                                           "inca_set_mixed_tail")
      call inca_check(inca_set_polynomials(inca_ctx, merge(1_c_int32_t, 0_c_int32_t, &
                      exact_polynomials)), "inca_set_polynomials")
      call inca_check(inca_set_point_cutoff(inca_ctx, point_cutoff), "inca_set_point_cutoff")
      call inca_check(inca_set_grouping(inca_ctx, merge(1_c_int32_t, 0_c_int32_t, &
                      engine_grouping)), "inca_set_grouping")
      write(*,*) "Starting GPU evaluation over primitive quartets..."

      t3 = wall()
      if (radial_integral .or. radial_plot .or. cubeintra) then
         call evaluate(ivec)
         call inca_check(inca_get_eval_counts(inca_ctx, evals_ref, evals_exec), &
                         "inca_get_eval_counts")
         if (radial_integral .and. learned_grid_used) call check_learned_grid(ivec)
         ! reductions on the main grid before any other grid replaces it
         if (radial_integral) then
            allocate(ipc(nquad), vpc(nquad))
            call inca_check(inca_integrate(inca_ctx, c_loc(rintegral), c_loc(vee), c_loc(ipc), &
                 c_loc(vpc)), "inca_integrate")
            u_max = maxval(dsqrt(sum(cent(:,1:nquad)**2, dim=1))) + 5.d0
            allocate(su(n_bins), sp(n_bins), sia(n_bins))
            write(*,*) "Starting Radial Smearing..."
            write(*,*) "Grid Points:", rrgrid, " Sigma:", 0.15d0
            call inca_check(inca_smearing(inca_ctx, u_max, int(n_bins, c_int32_t), 0.15d0, &
                 c_loc(su), c_loc(sp), c_loc(sia)), "inca_smearing")
         end if
         if (radial_plot) then
            allocate(rintra(max(nradi,1)))
            call inca_check(inca_radial_values(inca_ctx, c_loc(rintra)), "inca_radial_values")
         end if
      end if
      if (intracule_at_zero) then
         allocate(pts(3,1)); pts = ZERO
         call inca_check(inca_grid_points(inca_ctx, 1_c_int64_t, pts), "inca_grid_points")
         call evaluate_small(ivz)
         deallocate(pts)
      end if
      if (intracule_two_points) then
         allocate(pts(3,2))
         pts(:,1) = (/x_point1, y_point1, z_point1/)
         pts(:,2) = (/x_point2, y_point2, z_point2/)
         call inca_check(inca_grid_points(inca_ctx, 2_c_int64_t, pts), "inca_grid_points")
         call evaluate_small(iv2)
         deallocate(pts)
      end if
      t4 = wall()
      write(*,*) "Loop over primtives completed"

      np2 = int(nprim, c_int64_t)*(int(nprim, c_int64_t) + 1)/2
      refval = ((np2 + 1)*np2)/2
      if (counts(1) > refval) then
         write(*,*) "Full DM2prim read"
      else
         write(*,*) "Symmetry reduced DM2prim."
         if (.not. nosym) write(*,*) "Run with NOSYM option."
      end if
      write(*,*) ">> Intracule integrations completed. Writing output..."
      lim = thresh*(dble(nprim)*(dble(nprim) + ONE)*HALF)**(-ONE)

      write(iout,*) "=========================================================="
      write(iout,*) "               INCA INTRACULE CALCULATION                 "
      write(iout,*) "=========================================================="
      write(iout,*) "--- GRID AND ACCURACY SETTINGS ---"
      write(iout,*) "Original grid points:             ", maxgrid
      write(iout,*) "Symmetry reduced grid points:     ", rgrid
      write(iout,*) "Total surviving grid points:      ", rrgrid
      write(iout,*) "Quartets in the dm2p file:        ", counts(1)
      write(iout,*) "Quartets after 1st screening:     ", counts(2)
      write(iout,*) "Quartets actually computed:       ", counts(3)
      write(iout,'(A, ES12.4)') "Screening limit (lim):            ", lim
      ! the reference prints two thresholds it never sets (trsh1, trsh2);
      ! the line keeps its place, with zeros
      write(iout,*) "Thresholds used for DM2prim:      ", 0.d0, 0.d0
      write(iout,*) ""
      write(iout,*) "--- RESULTS ---"

      if (radial_integral) then
         write(iout,*) "Calculation Type: Radial Integral"
         write(iout,*) "Number of centres: ", nquad
         if (nquad > 1) write(iout,*) "Centres for radial integration:"
         do i = 1, nquad
            write(iout,*) i, ":::", cent(:,i)
         end do
         write(iout,*) "Weights for centres: ", Ps(:)
         write(iout,*) "Alpha parameter:     ", sfalpha(:)
         write(iout,*) "Gauss-Legendre nodes:", nradc(:)
         write(iout,*) "Gauss-Lebedev nodes: ", nangc(:)
         write(iout,*) "---------------------------------------------"
         write(iout,'(A, ES25.16)') 'TOTAL VALUE OF INTRACULE = ', rintegral
         write(iout,'(A, ES25.16)') 'TOTAL VALUE OF Vee       = ', vee
         npairs = nelec*(nelec - 1)/2
         write(iout,'(A, ES15.6)') 'Radial_integral error    = ', rintegral - dble(npairs)
         write(iout,*) "---------------------------------------------"
         write(iout,*) "-------LOCAL DECOMPOSITION (TFVC)-------"
         write(iout,'(A6,1X,A20,1X,A15,1X,A15)') "Center", "Coordinates", "Pairs(N_pairs)", &
              "Vee(Hartree)"
         do i = 1, nquad
            if (i == 1) then
               write(iout,'(I6,1X,3F7.3,1X,F15.8,1X,F15.8,1X,A)') i, cent(:,i), ipc(i), vpc(i), &
                    "(INTRA-ATOMIC)"
            else
               write(iout,'(I6,1X,3F7.3,1X,F15.8,1X,F15.8,1X,A)') i, cent(:,i), ipc(i), vpc(i), &
                    "(INTER/DISP)"
            end if
         end do
         write(iout,*) "---------------------------------------------"
      else if (radial_plot) then
         write(iout,*) "Calculation Type: Radial Plot, I(s) vs s"
         if (nradi > 0) write(iout,*) "Radius evaluated from ", radi(1), " to ", radi(nradi), " bohr"
         write(iout,*) "Data written to: ", trim(r_plot_name)
      else if (cubeintra) then
         write(iout,*) "Calculation Type: Vectorial Cubefile"
         write(iout,*) "Data written to: ", trim(cubeintraname)
      end if
      if (intracule_at_zero) write(iout,'(A, ES25.16)') 'INTRACULE AT ZERO = ', ivz(1)/TWO
      if (intracule_two_points) then
         write(iout,*) "Point 1: ", x_point1, y_point1, z_point1, " Intracule: ", iv2(1)
         write(iout,*) "Point 2: ", x_point2, y_point2, z_point2, " Intracule: ", iv2(2)
      end if

      write(iout,*) ""
      write(iout,*) "--- COMPUTATIONAL TIME ---"
      write(iout,'(A, F10.3, A)') "Total CPU time:             ", wall() - t1, " seconds"
      write(iout,'(A, F10.3, A)') "Grid generation time:       ", gridt, " seconds"
      write(iout,'(A, F10.3, A)') "Integration loop time:      ", t4 - t3, " seconds"
      write(iout,*) "=========================================================="
      ! engine lines after the reference's own blocks, so that a parser of the
      ! reference's layout reads the lines above unchanged
      write(iout,*) "--- GPU ENGINE ---"
      write(iout,*) "Polycoef fits zeroed (|C|>1e3):   ", counts(4)
      write(iout,'(A, 3F12.3, A)') "GPU kernels prep/eval/sum:  ", kern_ms, " ms"
      if (mixed_precision) then
         write(iout,*) "Engine precision: mixed (FP64 heavy quartets, FP32 tail ", mixed_tail, ")"
      else
         write(iout,*) "Engine precision: FP64"
      end if
      if (exact_polynomials) then
         write(iout,*) "Axis polynomials: exact expansion in rp"
      else
         write(iout,*) "Axis polynomials: Gauss-Hermite + Vandermonde fits (reference)"
      end if
      call inca_check(inca_get_groups(inca_ctx, ngroups, ngroups64, ncoef), "inca_get_groups")
      if (ngroups > 0) then
         write(iout,'(A, I0, A, I0, A)') " Evaluation: by shell-quartet groups (", ngroups, &
              " groups, ", ncoef, " coefficients)"
      else
         write(iout,*) "Evaluation: record by record"
      end if
      if (point_cutoff > 0.d0 .and. ngroups > 0) then
         write(iout,'(A, F6.1, A, I0, A, I0)') " Per-point cutoff: 2^-", point_cutoff, &
              " of each term's peak; group-point pairs executed ", evals_exec, &
              " for record-point evaluations ", evals_ref
      else if (point_cutoff > 0.d0) then
         write(iout,'(A, F6.1, A, I0, A, I0)') " Per-point cutoff: 2^-", point_cutoff, &
              " of each term's peak; evaluations executed ", evals_exec, " of ", evals_ref
      else
         write(iout,*) "Per-point cutoff: off (every term at every point)"
      end if
      ! an explicit format: list-directed output would wrap the long line
      if (len_trim(grid_note) > 0) write(iout,'(1X,A)') trim(grid_note)
      write(iout,*) "=========================================================="
      close(iout)

      if (radial_plot) then
         open(unit=16, file=r_plot_name, status='replace')
         write(16,*) "# --------------------------------------------------------"
         write(16,*) "# RADIAL INTRACULE DATA"
         write(16,*) "# Col 1: Distance s (bohr)"
         write(16,*) "# Col 2: I(s) (Radial Intracule)"
         write(16,*) "# Col 3: P(s) = 2 * pi * s^2 * I(s) (Probability Density)"
         write(16,*) "# Col 4: s * I(s)"
         write(16,*) "# --------------------------------------------------------"
         do i = 1, nradi
            write(16,'(F12.6, 3ES20.8)') radi(i), rintra(i), rintra(i)*TWO*pi*radi(i)**2, &
                 rintra(i)*TWO*pi*radi(i)
         end do
         close(16)
      end if

      if (radial_integral) call write_smearing(su, sp, sia)

      if (cubeintra) then
         open(unit=17, file=cubeintraname, status='replace')
         write(17,*) "CUBE FILE"
         write(17,*) "OUTER LOOP:X, MIDDLE LOOP:Y, INNER LOOP:Z"
         write(17,*) natoms, center_i(1), center_i(2), center_i(3)
         write(17,*) np_i(1), step_i(1), ZERO, ZERO
         write(17,*) np_i(2), ZERO, step_i(2), ZERO
         write(17,*) np_i(3), ZERO, ZERO, step_i(3)
         do i = 1, natoms
            write(17,*) an(i), chrg(i), cartes(i,1), cartes(i,2), cartes(i,3)
         end do
         do i = 1, rgrid
            write(17,*) ivec(i)
         end do
         close(17)
      end if
      write(*,*) ">> INCA job finished successfully. Output written to: ", trim(outname)
   end subroutine calculate_intracule

   !! The multicentre grid of the current centres (mod_intracule.f90:91-104).
   subroutine multicenter_grid()
      integer(c_int64_t) :: mg, rg, rrg
      call inca_check(inca_grid_multicenter(inca_ctx, int(nquad, c_int32_t), cent, &
           int(nradc, c_int32_t), int(nangc, c_int32_t), sfalpha, Ps, &
           merge(1_c_int32_t, 0_c_int32_t, nosym), mg, rg, rrg), "inca_grid_multicenter")
      maxgrid = int(mg); rgrid = int(rg); rrgrid = int(rrg)
   end subroutine multicenter_grid

   !! Sum-rule check of the learned autogrid.  Its integral is compared with
   !! the analytic integral of the prepared terms over all space (which every
   !! grid integral approximates).  The learned grid stands when its relative
   !! error is at most 1e-8, or no more than 1e-8 above the plain autogrid's
   !! (centercalc, evaluated on the same prepared quartets); otherwise the
   !! plain autogrid's centres, grid and Ivec replace it.
   subroutine check_learned_grid(iv)
      double precision, allocatable, target, intent(inout) :: iv(:)
      double precision, parameter :: tol = 1.d-8
      double precision, allocatable :: lc(:,:), lp(:), la(:)
      integer, allocatable :: lr(:), lg(:)
      double precision :: exact, err_l, err_p

      call inca_check(inca_get_analytic_integral(inca_ctx, exact), "inca_get_analytic_integral")
      err_l = grid_error()
      if (err_l <= tol) then
         write(grid_note,'(A,ES10.3,A,ES24.16)') "Learned autogrid: accepted, sum-rule error ", &
              err_l, "; analytic integral ", exact
         return
      end if
      lc = cent; lp = Ps; la = sfalpha; lr = nradc; lg = nangc
      write(*,*) "Learned grid sum-rule error", err_l, "; evaluating the plain autogrid"
      call centercalc()
      call multicenter_grid()
      call reevaluate()
      err_p = grid_error()
      if (err_l > err_p + tol) then
         write(grid_note,'(A,ES10.3,A,ES10.3,A,ES24.16)') "Learned autogrid: rejected, &
              &sum-rule error ", err_l, "; plain autogrid used, error ", err_p, &
              "; analytic integral ", exact
         return
      end if
      nquad = size(lp)
      deallocate(cent, Ps, sfalpha, nradc, nangc)
      cent = lc; Ps = lp; sfalpha = la; nradc = lr; nangc = lg
      call multicenter_grid()
      call reevaluate()
      write(grid_note,'(A,ES10.3,A,ES10.3,A,ES24.16)') "Learned autogrid: accepted, sum-rule &
           &error ", err_l, "; plain autogrid ", err_p, "; analytic integral ", exact
   contains
      double precision function grid_error()
         double precision, target :: tot, v
         double precision, allocatable, target :: ip(:), vp(:)
         allocate(ip(nquad), vp(nquad))
         call inca_check(inca_integrate(inca_ctx, c_loc(tot), c_loc(v), c_loc(ip), c_loc(vp)), &
              "inca_integrate")
         grid_error = abs(tot - exact)/abs(exact)
      end function grid_error
      subroutine reevaluate()
         integer :: n
         n = max(rrgrid, rgrid, 1)
         if (allocated(iv)) deallocate(iv)
         allocate(iv(n))
         iv = ZERO
         call inca_check(inca_evaluate_prepared(inca_ctx, merge(INCA_MIXED, INCA_FP64, &
              mixed_precision), c_loc(iv)), "inca_evaluate_prepared")
         call inca_check(inca_get_eval_counts(inca_ctx, evals_ref, evals_exec), &
              "inca_get_eval_counts")
      end subroutine reevaluate
   end subroutine check_learned_grid

   !! Evaluate Ivec on the engine's current grid and copy it to the host.
   subroutine evaluate(iv)
      double precision, allocatable, target, intent(inout) :: iv(:)
      integer :: n
      n = max(rrgrid, rgrid, 1)
      if (allocated(iv)) deallocate(iv)
      allocate(iv(n))
      iv = ZERO
      call run_engine(c_loc(iv))
   end subroutine evaluate

   subroutine evaluate_small(iv)
      double precision, target, intent(inout) :: iv(:)
      call run_engine(c_loc(iv))
   end subroutine evaluate_small

   subroutine run_engine(ptr)
      type(c_ptr), intent(in) :: ptr
      integer(c_int32_t) :: prec
      integer(c_int64_t), target :: cnt(4)
      prec = merge(INCA_MIXED, INCA_FP64, mixed_precision)
      call inca_check(inca_evaluate(inca_ctx, prec, thresh, ptr, c_loc(cnt)), "inca_evaluate")
      counts = cnt
      call inca_check(inca_get_timings(inca_ctx, kern_ms), "inca_get_timings")
   end subroutine run_engine

   !! <stem>_smearing.dat (mod_intracule.f90:535-550, 576-649)
   subroutine write_smearing(u, p, ia)
      double precision, intent(in) :: u(:), p(:), ia(:)
      character(len=200) :: smear_file
      integer :: dot_idx, j
      dot_idx = index(trim(outname), '.', back=.true.)
      if (dot_idx > 0) then
         smear_file = outname(1:dot_idx-1)//"_smearing.dat"
      else
         smear_file = trim(outname)//"_smearing.dat"
      end if
      open(unit=25, file=smear_file, status='replace')
      write(25,*) "#   u (bohr)       P(u) [Prob]      I_avg(u) [Density]"
      do j = 1, size(u)
         write(25,'(3ES16.8)') u(j), p(j), ia(j)
      end do
      close(25)
   end subroutine write_smearing

   double precision function wall()
      integer(8) :: c, r
      call system_clock(c, r)
      wall = dble(c)/dble(r)
   end function wall

end module intracule
