!! Automatic integration centres for the multicentre radial integral, with
!! the behaviour of the reference's src/mod_autogrid.f90:
!!  * centre 1 at the origin: population = sum of intra-atomic pairs
!!    Z(Z-1)/2, radius parameter = pair-weighted Bragg-Slater radius;
!!  * for every atom pair j<k (both heavy unless nohydro=.false.) the two
!!    intracule-space centres +-(R_j - R_k), population Z_j Z_k/2, radius
!!    (r_j + r_k)/2, with |z| < 0.5 snapped to 0;
!!  * centres closer than 0.1 bohr in every coordinate are merged
!!    (population-weighted radius, summed population);
!!  * every centre gets nrad radial and nang angular points.
module autogrid
   use geninfo
   use intrainfo
   use iso_c_binding
   use inca_hip, only: inca_mlp_create, inca_mlp_destroy, inca_mlp_load, inca_autogrid, &
                       inca_learned_autogrid, inca_mlp_error_text, c_string
   implicit none
contains
   !! Learned autogrid: the centres of centercalc with the Bragg-Slater radii
   !! of get_bs_radius replaced by the network's per-atom radii and one
   !! nrad/nang per centre from its classes (inca_learned_autogrid, network
   !! weights from learned_grid_weights).  calculate_intracule checks the
   !! resulting integral against the sum rule and falls back to centercalc.
   subroutine learned_centercalc()
      type(c_ptr) :: net
      integer(c_int) :: rc
      integer(c_int32_t) :: nq
      integer(c_int32_t), allocatable, target :: an32(:), tmn32(:), ra32(:), nr(:), na(:)
      double precision, allocatable, target :: c(:,:), p(:), a(:)
      integer :: i

      rc = inca_mlp_create(0_c_int32_t, 1_c_int64_t, net)
      if (rc /= 0) call fail("inca_mlp_create")
      rc = inca_mlp_load(net, c_string(learned_grid_weights))
      if (rc /= 0) call fail("inca_mlp_load")
      allocate(an32(natoms), tmn32(3*nprim), ra32(nprim))
      an32 = int(an(1:natoms), c_int32_t)
      tmn32 = reshape(int(TMN(1:nprim,1:3), c_int32_t), (/3*nprim/))
      ra32 = int(Ra(1:nprim), c_int32_t)
      rc = inca_learned_autogrid(net, int(natoms, c_int32_t), an32, cartes, int(nprim, c_int32_t), &
           Alpha, tmn32, ra32, merge(1_c_int32_t, 0_c_int32_t, nohydro), 0_c_int32_t, nq, &
           c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)
      if (rc /= 0 .and. rc /= -6) call fail("inca_learned_autogrid")
      allocate(c(3,nq), p(nq), a(nq), nr(nq), na(nq))
      rc = inca_learned_autogrid(net, int(natoms, c_int32_t), an32, cartes, int(nprim, c_int32_t), &
           Alpha, tmn32, ra32, merge(1_c_int32_t, 0_c_int32_t, nohydro), nq, nq, &
           c_loc(c), c_loc(p), c_loc(a), c_loc(nr), c_loc(na), c_null_ptr)
      if (rc /= 0) call fail("inca_learned_autogrid")
      rc = inca_mlp_destroy(net)
      nquad = int(nq)
      if (allocated(cent)) deallocate(cent, nradc, nangc, sfalpha, Ps)
      allocate(cent(3,nquad), nradc(nquad), nangc(nquad), sfalpha(nquad), Ps(nquad))
      cent = c; Ps = p; sfalpha = a; nradc = nr; nangc = na
      if (betaone) then
         sfalpha = 1.d0
         Ps = 1.d0
      end if
      do i = 1, nquad
         write(*,*) "Center ", i, " (learned) Weight(Pop)=", Ps(i), " Alpha=", sfalpha(i), &
                    " nrad=", nradc(i), " nang=", nangc(i)
      end do
      learned_grid_used = .true.
   contains
      subroutine fail(what)
         character(len=*), intent(in) :: what
         write(*,*) "ERROR in ", what, ": ", inca_mlp_error_text(net)
         stop 1
      end subroutine fail
   end subroutine learned_centercalc

   !! centercalc (src/mod_autogrid.f90:12-131) through the engine's
   !! inca_autogrid, the one implementation of the centre rules and of the
   !! Bragg-Slater table; every centre gets nrad radial and nang angular
   !! points.  betaone is applied by inca_autogrid.
   subroutine centercalc()
      integer(c_int) :: rc
      integer(c_int32_t) :: nq
      integer(c_int32_t), allocatable :: an32(:)
      double precision, allocatable, target :: c(:,:), p(:), a(:)
      integer :: i

      write(*,*) "Calculating centers for intracule function..."
      allocate(an32(natoms))
      an32 = int(an(1:natoms), c_int32_t)
      rc = inca_autogrid(int(natoms, c_int32_t), an32, cartes, &
           merge(1_c_int32_t, 0_c_int32_t, nohydro), merge(1_c_int32_t, 0_c_int32_t, betaone), &
           c_null_ptr, 0_c_int32_t, nq, c_null_ptr, c_null_ptr, c_null_ptr)
      if (rc /= 0 .and. rc /= -6) call stop_autogrid(rc)
      allocate(c(3,nq), p(nq), a(nq))
      rc = inca_autogrid(int(natoms, c_int32_t), an32, cartes, &
           merge(1_c_int32_t, 0_c_int32_t, nohydro), merge(1_c_int32_t, 0_c_int32_t, betaone), &
           c_null_ptr, nq, nq, c_loc(c), c_loc(p), c_loc(a))
      if (rc /= 0) call stop_autogrid(rc)
      nquad = int(nq)
      if (allocated(cent)) deallocate(cent, nradc, nangc, sfalpha, Ps)
      allocate(cent(3,nquad), nradc(nquad), nangc(nquad), sfalpha(nquad), Ps(nquad))
      cent = c; Ps = p; sfalpha = a
      do i = 1, nquad
         write(*,*) "Center ", i, " stored. Weight(Pop)=", Ps(i), " Alpha=", sfalpha(i)
      end do
      nradc = nrad
      nangc = nang
   contains
      subroutine stop_autogrid(code)
         integer(c_int), intent(in) :: code
         write(*,*) "ERROR in inca_autogrid: code ", code
         stop 1
      end subroutine stop_autogrid
   end subroutine centercalc
end module autogrid
