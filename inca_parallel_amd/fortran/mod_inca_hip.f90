!! ISO_C_BINDING interfaces to libinca_hip.so (include/inca_hip.h).
!! Logical flags cross as integer(c_int); strings as NUL-terminated
!! character(kind=c_char); arrays are passed by reference in the same
!! Fortran layouts the engine expects (r(3,G), tmn(nprim,3), cent(3,nquad)).
module inca_hip
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: INCA_FP64 = 0, INCA_MIXED = 1

   interface
      integer(c_int) function inca_create(device, ctx) bind(C, name="inca_create")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function inca_destroy(ctx) bind(C, name="inca_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function inca_last_error(ctx) bind(C, name="inca_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function inca_set_primitives(ctx, nprim, alpha, tmn, x, y, z, normalize) &
            bind(C, name="inca_set_primitives")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nprim, normalize
         real(c_double), intent(in) :: alpha(*), x(*), y(*), z(*)
         integer(c_int32_t), intent(in) :: tmn(*)
      end function
      integer(c_int) function inca_grid_multicenter(ctx, nquad, cent, nradc, nangc, sfalpha, &
            ps, nosym, maxgrid, rgrid, rrgrid) bind(C, name="inca_grid_multicenter")
         import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nquad, nosym
         real(c_double), intent(in) :: cent(*), sfalpha(*), ps(*)
         integer(c_int32_t), intent(in) :: nradc(*), nangc(*)
         integer(c_int64_t), intent(out) :: maxgrid, rgrid, rrgrid
      end function
      integer(c_int) function inca_grid_radial_plot(ctx, nblock, tart, step, nang, nosym, &
            nradi, ngrid) bind(C, name="inca_grid_radial_plot")
         import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nblock, nosym
         real(c_double), intent(in) :: tart(*), step(*)
         integer(c_int32_t), intent(in) :: nang(*)
         integer(c_int64_t), intent(out) :: nradi, ngrid
      end function
      integer(c_int) function inca_grid_cube(ctx, center, step, np, ngrid) &
            bind(C, name="inca_grid_cube")
         import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: center(3), step(3)
         integer(c_int32_t), intent(in) :: np(3)
         integer(c_int64_t), intent(out) :: ngrid
      end function
      integer(c_int) function inca_grid_points(ctx, n, r) bind(C, name="inca_grid_points")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: r(*)
      end function
      integer(c_int) function inca_get_grid(ctx, r, w, w_vee, center_map) &
            bind(C, name="inca_get_grid")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, r, w, w_vee, center_map
      end function
      integer(c_int) function inca_get_radial_layout(ctx, radi, smn) &
            bind(C, name="inca_get_radial_layout")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, radi, smn
      end function
      integer(c_int) function inca_load_dm2p(ctx, path, nrec, fmt) bind(C, name="inca_load_dm2p")
         import :: c_int, c_ptr, c_char, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int64_t), intent(out) :: nrec
         integer(c_int32_t), intent(out) :: fmt
      end function
      integer(c_int) function inca_evaluate(ctx, precision, thresh, ivec, counts) &
            bind(C, name="inca_evaluate")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx, ivec, counts
         integer(c_int32_t), value :: precision
         real(c_double), value :: thresh
      end function
      integer(c_int) function inca_set_mixed_tail(ctx, tail) bind(C, name="inca_set_mixed_tail")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: tail
      end function
      integer(c_int) function inca_set_polynomials(ctx, mode) bind(C, name="inca_set_polynomials")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: mode
      end function
      integer(c_int) function inca_set_point_cutoff(ctx, log2_rel) &
            bind(C, name="inca_set_point_cutoff")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: log2_rel
      end function
      integer(c_int) function inca_set_grouping(ctx, on) bind(C, name="inca_set_grouping")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: on
      end function
      integer(c_int) function inca_get_groups(ctx, ngroups, ngroups_fp64, ncoef) &
            bind(C, name="inca_get_groups")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: ngroups, ngroups_fp64, ncoef
      end function
      integer(c_int) function inca_get_eval_counts(ctx, reference, executed) &
            bind(C, name="inca_get_eval_counts")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: reference, executed
      end function
      integer(c_int) function inca_integrate(ctx, integral, vee, ipc, vpc) &
            bind(C, name="inca_integrate")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, integral, vee, ipc, vpc
      end function
      integer(c_int) function inca_radial_values(ctx, rintra) bind(C, name="inca_radial_values")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rintra
      end function
      integer(c_int) function inca_smearing(ctx, u_max, n_bins, sigma, u, p, iavg) &
            bind(C, name="inca_smearing")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: ctx
         real(c_double), value :: u_max, sigma
         integer(c_int32_t), value :: n_bins
         type(c_ptr), value :: u, p, iavg
      end function
      integer(c_int) function inca_get_timings(ctx, t) bind(C, name="inca_get_timings")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: t(3)
      end function
      integer(c_int) function inca_evaluate_prepared(ctx, precision, ivec) &
            bind(C, name="inca_evaluate_prepared")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx, ivec
         integer(c_int32_t), value :: precision
      end function
      integer(c_int) function inca_get_analytic_integral(ctx, total) &
            bind(C, name="inca_get_analytic_integral")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: total
      end function
      integer(c_int) function inca_mlp_create(device, seed, m) bind(C, name="inca_mlp_create")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         integer(c_int32_t), value :: device
         integer(c_int64_t), value :: seed
         type(c_ptr), intent(out) :: m
      end function
      integer(c_int) function inca_mlp_destroy(m) bind(C, name="inca_mlp_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: m
      end function
      type(c_ptr) function inca_mlp_last_error(m) bind(C, name="inca_mlp_last_error")
         import :: c_ptr
         type(c_ptr), value :: m
      end function
      integer(c_int) function inca_mlp_load(m, path) bind(C, name="inca_mlp_load")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function inca_onee(ctx, kind, idx, nmo, t, w, values) &
            bind(C, name="inca_onee")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx, t, w, values
         integer(c_int32_t), value :: kind, idx, nmo
      end function
      integer(c_int) function inca_autogrid(natoms, an, cartes, nohydro, betaone, radius, cap, &
            nquad, cent, ps, sfalpha) bind(C, name="inca_autogrid")
         import :: c_int, c_ptr, c_int32_t, c_double
         integer(c_int32_t), value :: natoms, nohydro, betaone, cap
         integer(c_int32_t), intent(in) :: an(*)
         real(c_double), intent(in) :: cartes(*)
         type(c_ptr), value :: radius, cent, ps, sfalpha
         integer(c_int32_t), intent(out) :: nquad
      end function
      integer(c_int) function inca_learned_autogrid(m, natoms, an, cartes, nprim, alpha, tmn, &
            prim_atom, nohydro, cap, nquad, cent, ps, sfalpha, nradc, nangc, radius) &
            bind(C, name="inca_learned_autogrid")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: m
         integer(c_int32_t), value :: natoms, nprim, nohydro, cap
         integer(c_int32_t), intent(in) :: an(*), tmn(*), prim_atom(*)
         real(c_double), intent(in) :: cartes(*), alpha(*)
         integer(c_int32_t), intent(out) :: nquad
         type(c_ptr), value :: cent, ps, sfalpha, nradc, nangc, radius
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

   type(c_ptr), save :: inca_ctx = c_null_ptr

contains

   !! Message of the last failing call, as a Fortran string.
   function inca_error_text(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = inca_last_error(ctx)
      n = int(c_strlen(p))
      allocate(character(len=n) :: msg)
      call c_f_pointer(p, chars, [n])
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function inca_error_text

   !! Stop the run on a failing engine call, in the reference's style
   !! (message on stdout, then `stop 1`).
   subroutine inca_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= 0) then
         write(*,*) "ERROR in ", trim(what), ": ", inca_error_text(inca_ctx)
         stop 1
      end if
   end subroutine inca_check

   !! Create the engine context on device 0 once per run.
   subroutine inca_open()
      integer(c_int) :: rc
      if (c_associated(inca_ctx)) return
      rc = inca_create(0_c_int, inca_ctx)
      if (rc /= 0) then
         write(*,*) "ERROR: cannot start the GPU engine: ", inca_error_text(c_null_ptr)
         stop 1
      end if
   end subroutine inca_open

   subroutine inca_close()
      integer(c_int) :: rc
      if (.not. c_associated(inca_ctx)) return
      rc = inca_destroy(inca_ctx)
      inca_ctx = c_null_ptr
   end subroutine inca_close

   !! Message of the last failing network call, as a Fortran string.
   function inca_mlp_error_text(m) result(msg)
      type(c_ptr), intent(in) :: m
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = inca_mlp_last_error(m)
      n = int(c_strlen(p))
      allocate(character(len=n) :: msg)
      call c_f_pointer(p, chars, [n])
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function inca_mlp_error_text

   !! NUL-terminated copy of a Fortran string for the C side.
   function c_string(s) result(cs)
      character(len=*), intent(in) :: s
      character(kind=c_char) :: cs(len_trim(s) + 1)
      integer :: i
      do i = 1, len_trim(s)
         cs(i) = s(i:i)
      end do
      cs(len_trim(s) + 1) = c_null_char
   end function c_string

end module inca_hip
