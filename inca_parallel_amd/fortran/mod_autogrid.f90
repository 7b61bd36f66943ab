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
   use inca_hip, only: inca_mlp_create, inca_mlp_destroy, inca_mlp_load, &
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

   subroutine centercalc()
      integer :: i, j, k, nqmax, sm, nq
      double precision, allocatable :: c1(:,:), pop(:), alf(:)
      double precision :: r_j, pairs, tot, wsum
      logical :: same

      nqmax = natoms*(natoms - 1) + 1
      allocate(c1(3,nqmax), pop(nqmax), alf(nqmax))
      c1 = 0.d0; pop = 0.d0; alf = 0.d0
      tot = 0.d0; wsum = 0.d0
      do i = 1, natoms
         r_j = get_bs_radius(an(i))
         pairs = an(i)*(an(i) - 1.d0)/2.d0
         tot = tot + pairs
         wsum = wsum + pairs*r_j
      end do
      pop(1) = tot
      if (tot > 1.d-15) then
         alf(1) = wsum/tot
      else
         alf(1) = get_bs_radius(1)
      end if
      write(*,*) "Calculating centers for intracule function..."
      sm = 1
      do j = 1, natoms - 1
         do k = j + 1, natoms
            if (nohydro .and. .not. (an(j) > 1 .and. an(k) > 1)) cycle
            sm = sm + 1
            c1(:,sm) = cartes(j,:) - cartes(k,:)
            pop(sm) = an(j)*an(k)*0.5d0
            alf(sm) = (get_bs_radius(an(j)) + get_bs_radius(an(k)))/2.d0
            if (abs(c1(3,sm)) < 0.5d0) c1(3,sm) = 0.d0
            c1(:,sm+1) = -c1(:,sm)
            pop(sm+1) = pop(sm)
            alf(sm+1) = alf(sm)
            sm = sm + 1
         end do
      end do
      nqmax = sm
      do i = 2, nqmax - 1
         if (sum(abs(c1(:,i))) == 0.d0) cycle
         do j = i + 1, nqmax
            same = all(abs(c1(:,i) - c1(:,j)) < 1.d-1)
            if (same) then
               alf(i) = (alf(i)*pop(i) + alf(j)*pop(j))/(pop(i) + pop(j))
               pop(i) = pop(i) + pop(j)
               c1(:,j) = 0.d0
               pop(j) = 0.d0
            end if
         end do
      end do
      nq = 0
      do i = 1, nqmax
         if (i == 1 .or. sum(abs(c1(:,i))) /= 0.d0) nq = nq + 1
      end do
      nquad = nq
      if (allocated(cent)) deallocate(cent, nradc, nangc, sfalpha, Ps)
      allocate(cent(3,nquad), nradc(nquad), nangc(nquad), sfalpha(nquad), Ps(nquad))
      nq = 0
      do i = 1, nqmax
         if (i == 1 .or. sum(abs(c1(:,i))) /= 0.d0) then
            nq = nq + 1
            cent(:,nq) = c1(:,i)
            Ps(nq) = pop(i)
            sfalpha(nq) = alf(i)
            write(*,*) "Center ", nq, " stored. Weight(Pop)=", Ps(nq), " Alpha=", sfalpha(nq)
         end if
      end do
      if (betaone) then
         sfalpha = 1.d0
         Ps = 1.d0
      end if
      nradc = nrad
      nangc = nang
      deallocate(c1, pop, alf)
   end subroutine centercalc

   !! Bragg-Slater radius in bohr: Slater's empirical radii (J. Chem. Phys.
   !! 41, 3199 (1964)) in Angstrom, H = 0.35 (Becke), He 0.35, Ne 0.38, and
   !! 1.00 for Ar and beyond, times 1.8897 (src/mod_autogrid.f90:142-185).
   double precision function get_bs_radius(iz)
      integer, intent(in) :: iz
      double precision, parameter :: r(36) = (/ 0.35d0, 0.35d0, 1.45d0, 1.05d0, 0.85d0, &
         0.70d0, 0.65d0, 0.60d0, 0.50d0, 0.38d0, 1.80d0, 1.50d0, 1.25d0, 1.10d0, 1.00d0, &
         1.00d0, 1.00d0, 1.00d0, 2.20d0, 1.80d0, 1.60d0, 1.40d0, 1.35d0, 1.40d0, 1.40d0, &
         1.40d0, 1.35d0, 1.35d0, 1.35d0, 1.35d0, 1.30d0, 1.25d0, 1.15d0, 1.15d0, 1.15d0, &
         1.00d0 /)
      if (iz >= 1 .and. iz <= 36) then
         get_bs_radius = r(iz)
      else
         get_bs_radius = 1.00d0
      end if
      get_bs_radius = get_bs_radius*1.8897d0
   end function get_bs_radius
end module autogrid
